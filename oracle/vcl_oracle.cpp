// vcl_oracle.cpp -- CPU restatement of the vcl differentiable-rendering hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: the GPU library in vcl_amd/ is
// compared against it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg may load it.  The product path never calls it.
//
// Every op is templated on the arithmetic type T:
//   T = float  : follows the reference kernels' float (and mixed float/double) arithmetic;
//                this is also the timed CPU baseline.
//   T = double : the FP64 yardstick used to state the error of both float paths.
// Coverage and triangle IDs are defined by the float software-raster spec below for both
// types (the double variant re-evaluates barycentrics/depth/DB of the same winners).
//
// Citations name files of the reference (236088/vcl, nvdiffrast/nvdiffrast/<file>:<line>).
// Parity status of each function is recorded in its comment:
//   pinned   -- checked in tests/ against closed forms, brute force, invariants or central
//               finite differences (listed per function).
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -pthread -shared).

#include "oracle_common.hpp"

int g_threads = 1;

extern "C" void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
extern "C" int orc_get_threads(void) { return g_threads; }

// ------------------------------------------------------------------------------------
// Project (project.cu:26-34 forward, :71-88 backward).  clip[v][i] = sum_j mat[4j+i]*[p,1]_j,
// mat column-major.  views > 1: one matrix and one clip slice per view.
// pinned: closed form and adjoint identity (tests/test_oracle.py::test_project_closed_form_and_adjoint).
template <class T>
static void project_fwd(int V, int views, const T* pos, const T* mats, T* clip) {
    for (int s = 0; s < views; s++) {
        const T* m = mats + 16 * s;
        for (int v = 0; v < V; v++) {
            T x = pos[3 * v], y = pos[3 * v + 1], z = pos[3 * v + 2];
            for (int i = 0; i < 4; i++)
                clip[((size_t)s * V + v) * 4 + i] = m[i] * x + m[4 + i] * y + m[8 + i] * z + m[12 + i];
        }
    }
}

// gpos[v] += M^T gclip[v], summed over views in view order.
template <class T>
static void project_bwd(int V, int views, const T* mats, const T* gclip, T* gpos) {
    for (int v = 0; v < V; v++) {
        for (int s = 0; s < views; s++) {
            const T* m = mats + 16 * s;
            const T* g = gclip + ((size_t)s * V + v) * 4;
            for (int i = 0; i < 4; i++) {
                gpos[3 * v] += m[i] * g[i];
                gpos[3 * v + 1] += m[4 + i] * g[i];
                gpos[3 * v + 2] += m[8 + i] * g[i];
            }
        }
    }
}

// ------------------------------------------------------------------------------------
// Software rasterizer spec (replaces the OpenGL raster of rasterize.cpp:91-203,257-282).
//
//  * Guard band / clipping.  A triangle whose three vertices satisfy w > 0, |x| <= GB*w and
//    |y| <= GB*w is rasterized as is.  Otherwise it is clipped (Sutherland-Hodgman, float)
//    against z + w >= 0, w >= WEPS and |x|,|y| <= GB*w, and the polygon is fanned from its
//    first vertex.  All pieces keep the original triangle ID.
//  * Snapping.  ndc = x / w; X = (ndc + 1) * (W/2) pixels; Xf = rint(X * 256) (8 sub-pixel bits).
//  * Coverage.  Pixel centre (px*256+128, py*256+128).  Edge functions in int64 with the
//    interior made positive; a centre exactly on an edge is covered only for a "top-left"
//    edge: A > 0 || (A == 0 && B < 0) where E = A*X + B*Y + C.  Zero-area pieces are skipped.
//    No culling (reference never enables GL_CULL_FACE).
//  * Depth.  z_ndc of the ORIGINAL triangle at the pixel centre pix (from double, as in
//    rasterize.cu:101-104) as Z(pix) / W(pix), the ratio of two affine functions whose
//    coefficients are fixed per triangle (depth_planes).  Fragments with w <= 0 (W and the
//    barycentric denominator D of opposite signs) or z_ndc outside [-1, 1] are dropped (GL clip
//    volume).  The visible fragment is the lexicographic minimum of (z_ndc, triangle id):
//    GL_LESS with draw-order ties (rasterize.cpp:252-254).  Barycentrics of the winner come from
//    the closed form of rasterize.cu:114-121; rast.z is the key's z_ndc.
//  * Output rast = (u, v, z_ndc, id + 1) (rasterize.cpp:175), background 0 (glClear).
//    DB = (du/dX, du/dY, dv/dX, dv/dY) per pixel, analytic (rasterize.cpp:100-164).
// pinned: brute force over all triangles per pixel (test_raster_matches_bruteforce_config1,
// test_raster_matches_bruteforce_soup_with_clipping); an independent float64 point-in-triangle
// and screen-plane depth test for "nearer triangle wins" and id ties
// (test_raster_nearer_triangle_wins_independent); watertight shared edges
// (test_raster_planar_grid_is_watertight, test_raster_closed_mesh_even_coverage); closed-form
// affine barycentrics (test_raster_affine_barycentrics_closed_form); DB vs central differences
// (test_raster_db_matches_finite_differences).

static const float RAST_GB = 16.0f;
static const float RAST_WEPS = 1e-6f;

struct RPiece { int X[3], Y[3]; int px0, px1, py0, py1; u32 id; i64 A[3], B[3], C[3]; int tl[3]; float zp[9]; };

// Depth planes of the ORIGINAL triangle.  With P_i = w_i*pix - p_i the closed-form numerators
// N_i = cross(P_j, P_k), (j, k) = (i+1, i+2) mod 3, are affine in pix:
//   N_i = (y_j w_k - w_j y_k) X + (w_j x_k - x_j w_k) Y + (x_j y_k - y_j x_k),
// so Z = sum z_i N_i, W = sum w_i N_i and D = sum N_i are affine too, and z/w = Z / W
// (the barycentric denominator cancels).  zp = (Az, Bz, Cz, Aw, Bw, Cw, Ad, Bd, Cd).
static void depth_planes(const float* c0, const float* c1, const float* c2, float* zp) {
    const float* c[3] = {c0, c1, c2};
    float a[3], b[3], d[3];
    for (int i = 0; i < 3; i++) {
        const float* j = c[(i + 1) % 3];
        const float* k = c[(i + 2) % 3];
        a[i] = j[1] * k[3] - j[3] * k[1];
        b[i] = j[3] * k[0] - j[0] * k[3];
        d[i] = j[0] * k[1] - j[1] * k[0];
    }
    zp[0] = (c0[2] * a[0] + c1[2] * a[1]) + c2[2] * a[2];
    zp[1] = (c0[2] * b[0] + c1[2] * b[1]) + c2[2] * b[2];
    zp[2] = (c0[2] * d[0] + c1[2] * d[1]) + c2[2] * d[2];
    zp[3] = (c0[3] * a[0] + c1[3] * a[1]) + c2[3] * a[2];
    zp[4] = (c0[3] * b[0] + c1[3] * b[1]) + c2[3] * b[2];
    zp[5] = (c0[3] * d[0] + c1[3] * d[1]) + c2[3] * d[2];
    zp[6] = (a[0] + a[1]) + a[2];
    zp[7] = (b[0] + b[1]) + b[2];
    zp[8] = (d[0] + d[1]) + d[2];
}

// Depth key of a fragment: false if the fragment is outside the clip volume (w <= 0 or
// z/w outside [-1, 1]), else z/w.
static inline bool frag_depth(const float* zp, float X, float Y, float& z) {
    float Zp = zp[0] * X + zp[1] * Y + zp[2];
    float Wp = zp[3] * X + zp[4] * Y + zp[5];
    float Dp = zp[6] * X + zp[7] * Y + zp[8];
    if (!((Wp > 0.0f && Dp > 0.0f) || (Wp < 0.0f && Dp < 0.0f))) return false;
    z = Zp / Wp;
    return z >= -1.0f && z <= 1.0f;
}

static inline u32 ord_f32(float z) {
    u32 b; std::memcpy(&b, &z, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

static bool setup_piece(const float* a, const float* b, const float* c, u32 id, int W, int H, RPiece& P) {
    const float* v[3] = {a, b, c};
    float halfW = 0.5f * (float)W, halfH = 0.5f * (float)H;
    for (int i = 0; i < 3; i++) {
        float nx = v[i][0] / v[i][3], ny = v[i][1] / v[i][3];
        float X = (nx + 1.0f) * halfW, Y = (ny + 1.0f) * halfH;
        P.X[i] = (int)std::rint(X * 256.0f);
        P.Y[i] = (int)std::rint(Y * 256.0f);
    }
    i64 area = (i64)(P.X[1] - P.X[0]) * (P.Y[2] - P.Y[0]) - (i64)(P.Y[1] - P.Y[0]) * (P.X[2] - P.X[0]);
    if (area == 0) return false;
    i64 s = area > 0 ? 1 : -1;
    for (int i = 0; i < 3; i++) {
        int j = (i + 1) % 3;
        i64 dx = P.X[j] - P.X[i], dy = P.Y[j] - P.Y[i];
        // E = s*(dx*(Y - Yi) - dy*(X - Xi)) = A*X + B*Y + C
        P.A[i] = -s * dy; P.B[i] = s * dx;
        P.C[i] = s * (dy * P.X[i] - dx * P.Y[i]);
        P.tl[i] = (P.A[i] > 0 || (P.A[i] == 0 && P.B[i] < 0)) ? 1 : 0;
    }
    int xmin = std::min(P.X[0], std::min(P.X[1], P.X[2])), xmax = std::max(P.X[0], std::max(P.X[1], P.X[2]));
    int ymin = std::min(P.Y[0], std::min(P.Y[1], P.Y[2])), ymax = std::max(P.Y[0], std::max(P.Y[1], P.Y[2]));
    P.px0 = std::max(0, (xmin - 128 + 255) >> 8); P.px1 = std::min(W - 1, (xmax - 128) >> 8);
    P.py0 = std::max(0, (ymin - 128 + 255) >> 8); P.py1 = std::min(H - 1, (ymax - 128) >> 8);
    P.id = id;
    return P.px0 <= P.px1 && P.py0 <= P.py1;
}

static inline bool piece_covers(const RPiece& P, int px, int py) {
    i64 X = (i64)px * 256 + 128, Y = (i64)py * 256 + 128;
    for (int i = 0; i < 3; i++) {
        i64 e = P.A[i] * X + P.B[i] * Y + P.C[i];
        if (!(e > 0 || (e == 0 && P.tl[i]))) return false;
    }
    return true;
}

static inline bool vtx_fast(const float* v) {
    return v[3] > 0.0f && std::fabs(v[0]) <= RAST_GB * v[3] && std::fabs(v[1]) <= RAST_GB * v[3];
}

// Clip polygon against plane dot(pl, v) >= 0 (float Sutherland-Hodgman).
static int clip_poly(const float in[][4], int n, float out[][4], const float pl[5]) {
    int m = 0;
    for (int i = 0; i < n; i++) {
        const float* a = in[i]; const float* b = in[(i + 1) % n];
        float da = pl[0] * a[0] + pl[1] * a[1] + pl[2] * a[2] + pl[3] * a[3] + pl[4];
        float db = pl[0] * b[0] + pl[1] * b[1] + pl[2] * b[2] + pl[3] * b[3] + pl[4];
        if (da >= 0.0f) { for (int k = 0; k < 4; k++) out[m][k] = a[k]; m++; }
        if ((da >= 0.0f) != (db >= 0.0f)) {
            float t = da / (da - db);
            for (int k = 0; k < 4; k++) out[m][k] = a[k] + t * (b[k] - a[k]);
            m++;
        }
    }
    return m;
}

// Produce the raster pieces of triangle `id`; returns count (<= 8).
static int make_pieces(const float* v0, const float* v1, const float* v2, u32 id, int W, int H, RPiece* out) {
    float zp[9];
    depth_planes(v0, v1, v2, zp);
    if (vtx_fast(v0) && vtx_fast(v1) && vtx_fast(v2)) {
        if (!setup_piece(v0, v1, v2, id, W, H, out[0])) return 0;
        std::memcpy(out[0].zp, zp, sizeof(zp));
        return 1;
    }
    static const float planes[6][5] = {
        {0, 0, 1, 1, 0}, {0, 0, 0, 1, -RAST_WEPS},
        {-1, 0, 0, RAST_GB, 0}, {1, 0, 0, RAST_GB, 0}, {0, -1, 0, RAST_GB, 0}, {0, 1, 0, RAST_GB, 0}};
    float A[12][4], B[12][4];
    for (int k = 0; k < 4; k++) { A[0][k] = v0[k]; A[1][k] = v1[k]; A[2][k] = v2[k]; }
    int n = 3;
    for (int p = 0; p < 6 && n > 0; p++) {
        n = clip_poly(A, n, B, planes[p]);
        for (int i = 0; i < n; i++) for (int k = 0; k < 4; k++) A[i][k] = B[i][k];
    }
    int cnt = 0;
    for (int i = 1; i + 1 < n; i++)
        if (setup_piece(A[0], A[i], A[i + 1], id, W, H, out[cnt])) { std::memcpy(out[cnt].zp, zp, sizeof(zp)); cnt++; }
    return cnt;
}

// Barycentrics and depth of the original triangle at pixel (px,py), float, fixed op order
// (closed form of rasterize.cu:101-121; pix as there: double then rounded).
struct BaryF { float u, v, z, wc; float P0x, P0y, P1x, P1y, P2x, P2y, D; float pixx, pixy; };
static inline void bary_f32(const float* c0, const float* c1, const float* c2, int px, int py, float xs, float ys, BaryF& r) {
    float pixx = (float)((px + 0.5) * (double)xs - 1.0);
    float pixy = (float)((py + 0.5) * (double)ys - 1.0);
    float P0x = c0[3] * pixx - c0[0], P0y = c0[3] * pixy - c0[1];
    float P1x = c1[3] * pixx - c1[0], P1y = c1[3] * pixy - c1[1];
    float P2x = c2[3] * pixx - c2[0], P2y = c2[3] * pixy - c2[1];
    float E0x = P0x - P2x, E0y = P0y - P2y, E1x = P1x - P2x, E1y = P1y - P2y;
    float D = E0x * E1y - E0y * E1x;
    float u = (P1x * P2y - P1y * P2x) / D;
    float v = (P2x * P0y - P2y * P0x) / D;
    float b2 = 1.0f - u - v;
    float zc = c0[2] * u + c1[2] * v + c2[2] * b2;
    float wc = c0[3] * u + c1[3] * v + c2[3] * b2;
    r.u = u; r.v = v; r.wc = wc; r.z = zc / wc;
    r.P0x = P0x; r.P0y = P0y; r.P1x = P1x; r.P1y = P1y; r.P2x = P2x; r.P2y = P2y; r.D = D;
    r.pixx = pixx; r.pixy = pixy;
}

static inline bool frag_ok(const BaryF& b) { return b.wc > 0.0f && b.z >= -1.0f && b.z <= 1.0f; }

// Generic-T barycentrics, depth and DB (for the double yardstick and the float outputs).
template <class T>
static void bary_db(const T* c0, const T* c1, const T* c2, int px, int py, T xs, T ys, T* rast4, T* db4) {
    T pixx = (T)((px + 0.5) * (double)xs - 1.0);
    T pixy = (T)((py + 0.5) * (double)ys - 1.0);
    T w0 = c0[3], w1 = c1[3], w2 = c2[3];
    T P0x = w0 * pixx - c0[0], P0y = w0 * pixy - c0[1];
    T P1x = w1 * pixx - c1[0], P1y = w1 * pixy - c1[1];
    T P2x = w2 * pixx - c2[0], P2y = w2 * pixy - c2[1];
    T E0x = P0x - P2x, E0y = P0y - P2y, E1x = P1x - P2x, E1y = P1y - P2y;
    T D = E0x * E1y - E0y * E1x;
    T u = (P1x * P2y - P1y * P2x) / D;
    T v = (P2x * P0y - P2y * P0x) / D;
    T b2 = T(1) - u - v;
    T zc = c0[2] * u + c1[2] * v + c2[2] * b2;
    T wc = w0 * u + w1 * v + w2 * b2;
    rast4[0] = u; rast4[1] = v; rast4[2] = zc / wc;
    if (db4) {
        T dNuX = xs * (w1 * P2y - w2 * P1y), dNuY = ys * (w2 * P1x - w1 * P2x);
        T dNvX = xs * (w2 * P0y - w0 * P2y), dNvY = ys * (w0 * P2x - w2 * P0x);
        T dDX = xs * ((w0 - w2) * E1y - (w1 - w2) * E0y);
        T dDY = ys * ((w1 - w2) * E0x - (w0 - w2) * E1x);
        db4[0] = (dNuX - u * dDX) / D;
        db4[1] = (dNuY - u * dDY) / D;
        db4[2] = (dNvX - v * dDX) / D;
        db4[3] = (dNvY - v * dDY) / D;
    }
}

// z/w of the winner at the pixel centre by the depth planes (same ops as the key for float).
template <class T>
static T depth_at(const T* c0, const T* c1, const T* c2, int px, int py, T xs, T ys) {
    const T* c[3] = {c0, c1, c2};
    T a[3], b[3], d[3];
    for (int i = 0; i < 3; i++) {
        const T* j = c[(i + 1) % 3];
        const T* k = c[(i + 2) % 3];
        a[i] = j[1] * k[3] - j[3] * k[1];
        b[i] = j[3] * k[0] - j[0] * k[3];
        d[i] = j[0] * k[1] - j[1] * k[0];
    }
    T X = (T)((px + 0.5) * (double)xs - 1.0), Y = (T)((py + 0.5) * (double)ys - 1.0);
    T Az = (c0[2] * a[0] + c1[2] * a[1]) + c2[2] * a[2], Bz = (c0[2] * b[0] + c1[2] * b[1]) + c2[2] * b[2];
    T Cz = (c0[2] * d[0] + c1[2] * d[1]) + c2[2] * d[2];
    T Aw = (c0[3] * a[0] + c1[3] * a[1]) + c2[3] * a[2], Bw = (c0[3] * b[0] + c1[3] * b[1]) + c2[3] * b[2];
    T Cw = (c0[3] * d[0] + c1[3] * d[1]) + c2[3] * d[2];
    return (Az * X + Bz * Y + Cz) / (Aw * X + Bw * Y + Cw);
}

// Winner keys per pixel of one view: (ord(z) << 32) | id, ~0 = empty.
static void raster_keys(int W, int H, const float* clip, int T_, const u32* idx, std::vector<u64>& key) {
    key.assign((size_t)W * H, ~0ull);
    float xs = 2.0f / (float)W, ys = 2.0f / (float)H;
    // pieces for all triangles
    std::vector<RPiece> pieces; pieces.reserve(T_);
    RPiece tmp[8];
    for (int t = 0; t < T_; t++) {
        const float* c0 = clip + 4 * (size_t)idx[3 * t];
        const float* c1 = clip + 4 * (size_t)idx[3 * t + 1];
        const float* c2 = clip + 4 * (size_t)idx[3 * t + 2];
        int n = make_pieces(c0, c1, c2, (u32)t, W, H, tmp);
        for (int i = 0; i < n; i++) pieces.push_back(tmp[i]);
    }
    // rows split over threads; each thread walks all pieces but writes only its rows
    parallel_for(H, [&](int y0, int y1, int) {
        for (const RPiece& P : pieces) {
            int ya = std::max(P.py0, y0), yb = std::min(P.py1, y1 - 1);
            if (ya > yb) continue;
            for (int py = ya; py <= yb; py++) {
                float Y = (float)((py + 0.5) * (double)ys - 1.0);
                for (int px = P.px0; px <= P.px1; px++) {
                    if (!piece_covers(P, px, py)) continue;
                    float X = (float)((px + 0.5) * (double)xs - 1.0), z;
                    if (!frag_depth(P.zp, X, Y, z)) continue;
                    u64 k = ((u64)ord_f32(z) << 32) | P.id;
                    u64& dst = key[(size_t)py * W + px];
                    if (k < dst) dst = k;
                }
            }
        }
    });
}

template <class T>
static void raster_resolve(int W, int H, const T* clip, const u32* idx, const std::vector<u64>& key, T* rast, T* db) {
    T xs = (T)(2.0f / (float)W), ys = (T)(2.0f / (float)H);
    parallel_for(H, [&](int y0, int y1, int) {
        for (int py = y0; py < y1; py++)
            for (int px = 0; px < W; px++) {
                size_t p = (size_t)py * W + px;
                T* r = rast + 4 * p; T* d = db ? db + 4 * p : nullptr;
                u64 k = key[p];
                if (k == ~0ull) { r[0] = r[1] = r[2] = r[3] = 0; if (d) d[0] = d[1] = d[2] = d[3] = 0; continue; }
                u32 id = (u32)(k & 0xffffffffu);
                const T* c0 = clip + 4 * (size_t)idx[3 * id];
                const T* c1 = clip + 4 * (size_t)idx[3 * id + 1];
                const T* c2 = clip + 4 * (size_t)idx[3 * id + 2];
                bary_db<T>(c0, c1, c2, px, py, xs, ys, r, d);
                r[2] = depth_at<T>(c0, c1, c2, px, py, xs, ys);
                r[3] = (T)(float)(id + 1);
            }
    });
}

// Brute force: per pixel, test every piece of every triangle.  Test oracle for the oracle.
extern "C" void orc_raster_brute(int W, int H, const float* clip, int T_, const u32* idx, u32* ids /*P, 0 = empty, else id+1*/) {
    float xs = 2.0f / (float)W, ys = 2.0f / (float)H;
    std::vector<RPiece> pieces; RPiece tmp[8];
    for (int t = 0; t < T_; t++) {
        int n = make_pieces(clip + 4 * (size_t)idx[3 * t], clip + 4 * (size_t)idx[3 * t + 1], clip + 4 * (size_t)idx[3 * t + 2], (u32)t, W, H, tmp);
        for (int i = 0; i < n; i++) pieces.push_back(tmp[i]);
    }
    for (int py = 0; py < H; py++)
        for (int px = 0; px < W; px++) {
            u64 best = ~0ull;
            float X = (float)((px + 0.5) * (double)xs - 1.0), Y = (float)((py + 0.5) * (double)ys - 1.0);
            for (const RPiece& P : pieces) {
                if (!piece_covers(P, px, py)) continue;
                float z;
                if (!frag_depth(P.zp, X, Y, z)) continue;
                u64 k = ((u64)ord_f32(z) << 32) | P.id;
                if (k < best) best = k;
            }
            ids[(size_t)py * W + px] = best == ~0ull ? 0u : (u32)(best & 0xffffffffu) + 1u;
        }
}

// Coverage count per pixel over all pieces (no depth test): used by watertightness tests.
extern "C" void orc_raster_coverage_count(int W, int H, const float* clip, int T_, const u32* idx, u32* count) {
    std::memset(count, 0, sizeof(u32) * (size_t)W * H);
    RPiece tmp[8];
    for (int t = 0; t < T_; t++) {
        int n = make_pieces(clip + 4 * (size_t)idx[3 * t], clip + 4 * (size_t)idx[3 * t + 1], clip + 4 * (size_t)idx[3 * t + 2], (u32)t, W, H, tmp);
        for (int i = 0; i < n; i++)
            for (int py = tmp[i].py0; py <= tmp[i].py1; py++)
                for (int px = tmp[i].px0; px <= tmp[i].px1; px++)
                    if (piece_covers(tmp[i], px, py)) count[(size_t)py * W + px]++;
    }
}

// cov_clip (optional): float clip coordinates that decide coverage and visibility instead of
// float(clip).  The FP64 yardstick passes the float path's own projection, so it resolves the
// same winners as the float path (a double projection rounded to float can differ from the
// float projection in the last bit and flip ties) and evaluates everything continuous in double.
template <class T>
static void raster_fwd(int W, int H, int views, int V, const T* clip, int T_, const u32* idx, T* rast, T* db, const float* cov_clip = nullptr) {
    std::vector<float> cf((size_t)V * 4);
    std::vector<u64> key;
    for (int s = 0; s < views; s++) {
        const T* cs = clip + (size_t)s * V * 4;
        if (cov_clip) std::memcpy(cf.data(), cov_clip + (size_t)s * V * 4, sizeof(float) * cf.size());
        else for (size_t i = 0; i < cf.size(); i++) cf[i] = (float)cs[i];
        raster_keys(W, H, cf.data(), T_, idx, key);
        raster_resolve<T>(W, H, cs, idx, key, rast + (size_t)s * W * H * 4, db ? db + (size_t)s * W * H * 4 : nullptr);
    }
}

// RasterizeBackwardKernel (rasterize.cu:84-141): (dL/du, dL/dv) -> clip x, y, w of the three
// vertices.  grast is P x 2 (float2 per pixel, rasterize.cu:94).  z and DB grads ignored
// (rasterize.cu:96-98,134-136).
// pinned: central differences of the closed-form u,v w.r.t. clip coords
// (test_raster_bwd_matches_finite_differences).
template <class T>
static void raster_bwd(int W, int H, int views, int V, const T* clip, const u32* idx, const T* rast, const T* grast, T* gclip) {
    T xs = (T)(2.0f / (float)W), ys = (T)(2.0f / (float)H);
    Scatter<T> sc(gclip, (size_t)views * V * 4);
    parallel_for(views * H, [&](int r0, int r1, int th) {
        T* g = sc.buf(th);
        for (int row = r0; row < r1; row++) {
            int s = row / H, py = row % H;
            const T* cl = clip + (size_t)s * V * 4;
            T* gs = g + (size_t)s * V * 4;
            for (int px = 0; px < W; px++) {
                size_t p = ((size_t)s * H + py) * W + px;
                const T* r = rast + 4 * p;
                int id = (int)r[3] - 1;
                if (id < 0) continue;
                T gu = grast[2 * p], gv = grast[2 * p + 1];
                if (gu == T(0) && gv == T(0)) continue;
                T pixx = (T)((px + 0.5) * (double)xs - 1.0), pixy = (T)((py + 0.5) * (double)ys - 1.0);
                u32 i0 = idx[3 * id], i1 = idx[3 * id + 1], i2 = idx[3 * id + 2];
                const T *c0 = cl + 4 * (size_t)i0, *c1 = cl + 4 * (size_t)i1, *c2 = cl + 4 * (size_t)i2;
                T p0x = c0[3] * pixx - c0[0], p0y = c0[3] * pixy - c0[1];
                T p1x = c1[3] * pixx - c1[0], p1y = c1[3] * pixy - c1[1];
                T p2x = c2[3] * pixx - c2[0], p2y = c2[3] * pixy - c2[1];
                T e0x = p0x - p2x, e0y = p0y - p2y, e1x = p1x - p2x, e1y = p1y - p2y;
                T a = e0x * e1y - e0y * e1x;
                T ia = T(1) / (a + (a > 0 ? T(1e-6f) : T(-1e-6f)));
                T kuv = r[0] * gu + r[1] * gv;
                T gx0 = (gv * p2y + kuv * e1y) * ia;
                T gy0 = (-gv * p2x - kuv * e1x) * ia;
                T gw0 = -gx0 * pixx - gy0 * pixy;
                T gx1 = (-gu * p2y - kuv * e0y) * ia;
                T gy1 = (gu * p2x + kuv * e0x) * ia;
                T gw1 = -gx1 * pixx - gy1 * pixy;
                T gx2 = ((-gv * p0y + gu * p1y) - kuv * (-p0y + p1y)) * ia;
                T gy2 = ((gv * p0x - gu * p1x) - kuv * (p0x - p1x)) * ia;
                T gw2 = -gx2 * pixx - gy2 * pixy;
                T* q;
                q = gs + 4 * (size_t)i0; q[0] += gx0; q[1] += gy0; q[3] += gw0;
                q = gs + 4 * (size_t)i1; q[0] += gx1; q[1] += gy1; q[3] += gw1;
                q = gs + 4 * (size_t)i2; q[0] += gx2; q[1] += gy2; q[3] += gw2;
            }
        }
    });
    sc.merge();
}

// ------------------------------------------------------------------------------------
// Interpolate (interpolate.cu:20-49 forward, :114-143 backward).
// out = a0*u + a1*v + a2*(1.0-u-v) with the double literal of interpolate.cu:38 (A-14);
// DA_i = ((a0-a2)*db.x + (a1-a2)*db.z, (a0-a2)*db.y + (a1-a2)*db.w).  Background stays 0.
// pinned: linear attributes reproduce exactly, adjoint identity of the backward
// (test_interpolate_linear_attributes_and_adjoint).
template <class T>
static void interp_fwd(int W, int H, int views, int D, const T* rast, const T* db, const T* attr, const u32* idx, T* out, T* outDA) {
    parallel_for(views * H, [&](int r0, int r1, int) {
        for (int row = r0; row < r1; row++)
            for (int px = 0; px < W; px++) {
                size_t p = (size_t)row * W + px;
                const T* r = rast + 4 * p;
                int id = (int)r[3] - 1;
                if (id < 0) {
                    for (int i = 0; i < D; i++) { out[p * D + i] = 0; if (outDA) { outDA[2 * (p * D + i)] = 0; outDA[2 * (p * D + i) + 1] = 0; } }
                    continue;
                }
                u32 i0 = idx[3 * id], i1 = idx[3 * id + 1], i2 = idx[3 * id + 2];
                for (int i = 0; i < D; i++) {
                    T a0 = attr[(size_t)i0 * D + i], a1 = attr[(size_t)i1 * D + i], a2 = attr[(size_t)i2 * D + i];
                    out[p * D + i] = (T)(a0 * r[0] + a1 * r[1] + a2 * (1.0 - r[0] - r[1]));
                    if (outDA) {
                        const T* b = db + 4 * p;
                        T dadu = a0 - a2, dadv = a1 - a2;
                        outDA[2 * (p * D + i)] = dadu * b[0] + dadv * b[2];
                        outDA[2 * (p * D + i) + 1] = dadu * b[1] + dadv * b[3];
                    }
                }
            }
    });
}

// gattr += (atomics of interpolate.cu:133-135); grast[p] = (gu, gv) stored (interpolate.cu:141).
template <class T>
static void interp_bwd(int W, int H, int views, int D, const T* rast, const T* attr, const u32* idx, const T* gout, T* gattr, T* grast, int V) {
    Scatter<T> sc(gattr, (size_t)V * D);
    parallel_for(views * H, [&](int r0, int r1, int th) {
        T* ga = sc.buf(th);
        for (int row = r0; row < r1; row++)
            for (int px = 0; px < W; px++) {
                size_t p = (size_t)row * W + px;
                const T* r = rast + 4 * p;
                int id = (int)r[3] - 1;
                if (id < 0) continue;
                u32 i0 = idx[3 * id], i1 = idx[3 * id + 1], i2 = idx[3 * id + 2];
                T gu = 0, gv = 0;
                for (int i = 0; i < D; i++) {
                    T g = gout[p * D + i];
                    if (ga) {
                        ga[(size_t)i0 * D + i] += g * r[0];
                        ga[(size_t)i1 * D + i] += g * r[1];
                        ga[(size_t)i2 * D + i] += (T)(g * (1.0 - r[0] - r[1]));
                    }
                    gu += g * (attr[(size_t)i0 * D + i] - attr[(size_t)i2 * D + i]);
                    gv += g * (attr[(size_t)i1 * D + i] - attr[(size_t)i2 * D + i]);
                }
                if (grast) { grast[2 * p] = gu; grast[2 * p + 1] = gv; }
            }
    });
    sc.merge();
}

// ------------------------------------------------------------------------------------
// Mip texture (texturemap.cu:22-95 forward, :160-284 backward; buffer.cu:379-484,615-648).
// Levels are stored back to back: level l is (texW>>l) x (texH>>l) x C, row-major HWC.
static size_t level_offset(int tw, int th, int C, int l) {
    size_t off = 0;
    for (int k = 0; k < l; k++) off += (size_t)(tw >> k) * (th >> k) * C;
    return off;
}

// Mip level count of Texture::init (buffer.cu:380-385): min(LSB(w|h)+1, 16, requested), >= 1.
extern "C" int orc_mip_levels(int w, int h, int requested) {
    unsigned v = (unsigned)(w | h);
    int lsb = 0;
    while (lsb < 31 && !((v >> lsb) & 1u)) lsb++;
    int maxl = std::min(lsb + 1, 16);
    return requested < 1 ? 1 : std::min(maxl, requested);
}

// calculateLevel (texturemap.cu:22-47).
template <class T>
static void calc_level(const T* uvDA, int tw, int th, int L, int& l0, int& l1, T& f, T& level_out) {
    T dsdx = uvDA[0] * tw, dsdy = uvDA[1] * tw, dtdx = uvDA[2] * th, dtdy = uvDA[3] * th;
    T s2 = dsdx * dsdx + dsdy * dsdy, t2 = dtdx * dtdx + dtdy * dtdy;
    T b = T(0.5f) * (s2 + t2);
    // c^2 = b^2 - a^2 = (b - a)(b + a) with each factor a sum of squares: the same value as
    // texturemap.cu:41 without its cancellation for near-conformal footprints (A-20)
    T pm = dsdx - dtdy, qm = dsdy + dtdx, pp = dsdx + dtdy, qp = dsdy - dtdx;
    T c2 = T(0.25f) * ((pm * pm + qm * qm) * (pp * pp + qp * qp));
    T c = std::sqrt(c2);
    T level = T(0.5f) * std::log2(b + c);
    l0 = level <= 0 ? 0 : (T)(L - 2) <= level ? L - 2 : (int)std::floor(level);
    l1 = level <= 1 ? 1 : (T)(L - 1) <= level ? L - 1 : (int)std::floor(level) + 1;
    f = level <= 0 ? T(0) : (T)(L - 1) <= level ? T(1) : level - std::floor(level);
    level_out = level;
}

// indexFetch (texturemap.cu:49-65), including the no-half-texel-offset taps (A-1).
template <class T>
static void index_fetch(int tw, int th, int level, T uvx, T uvy, int idx4[4], T& fx, T& fy) {
    int sx = tw >> level, sy = th >> level;
    T tx = uvx * (T)sx, ty = uvy * (T)sy;
    int u0 = tx < 0 ? 0 : tx > sx - 1 ? sx - 1 : (int)tx;
    int u1 = tx < 1 ? 0 : tx > sx - 2 ? sx - 1 : (int)tx + 1;
    int v0 = ty < 0 ? 0 : ty > sy - 1 ? sy - 1 : (int)ty;
    int v1 = ty < 1 ? 0 : ty > sy - 2 ? sy - 1 : (int)ty + 1;
    idx4[0] = v0 * sx + u0; idx4[1] = v0 * sx + u1; idx4[2] = v1 * sx + u0; idx4[3] = v1 * sx + u1;
    fx = tx < 0 ? T(0) : (T)sx < tx ? T(1) : tx - std::floor(tx);
    fy = ty < 0 ? T(0) : (T)sy < ty ? T(1) : ty - std::floor(ty);
}

template <class T> static inline T lerp_(T a, T b, T t) { return a + t * (b - a); }
template <class T> static inline T bilerp_(T a, T b, T c, T d, T tx, T ty) { return lerp_(lerp_(a, b, tx), lerp_(c, d, tx), ty); }

// pinned: affine textures, whose box mips stay affine, give a closed-form trilinear sample at
// any uv and isotropic footprint (test_texture_affine_trilinear_closed_form); constant texture
// invariance and an exact 2^k footprint -> level k (test_texture_constant_and_exact_level).
template <class T>
static void tex_fwd(int W, int H, int views, int tw, int th, int C, int L, const T* levels, const T* rast, const T* uv, const T* uvDA, T* out) {
    std::vector<size_t> off(L);
    for (int l = 0; l < L; l++) off[l] = level_offset(tw, th, C, l);
    parallel_for(views * H, [&](int r0, int r1, int) {
        for (int row = r0; row < r1; row++)
            for (int px = 0; px < W; px++) {
                size_t p = (size_t)row * W + px;
                T* o = out + p * C;
                if (rast[4 * p + 3] < T(1)) { for (int c = 0; c < C; c++) o[c] = 0; continue; }
                T ux = uv[2 * p], uy = uv[2 * p + 1];
                if (L > 1) {
                    int l0, l1; T f, lev;
                    calc_level<T>(uvDA + 4 * p, tw, th, L, l0, l1, f, lev);
                    int i0[4], i1[4]; T f0x, f0y, f1x, f1y;
                    index_fetch<T>(tw, th, l0, ux, uy, i0, f0x, f0y);
                    index_fetch<T>(tw, th, l1, ux, uy, i1, f1x, f1y);
                    const T* t0 = levels + off[l0]; const T* t1 = levels + off[l1];
                    for (int c = 0; c < C; c++) {
                        T v = bilerp_(t0[i0[0] * C + c], t0[i0[1] * C + c], t0[i0[2] * C + c], t0[i0[3] * C + c], f0x, f0y);
                        if (f > 0) {
                            T v1 = bilerp_(t1[i1[0] * C + c], t1[i1[1] * C + c], t1[i1[2] * C + c], t1[i1[3] * C + c], f1x, f1y);
                            v = lerp_(v, v1, f);
                        }
                        o[c] = v;
                    }
                } else {
                    int i0[4]; T fx, fy;
                    index_fetch<T>(tw, th, 0, ux, uy, i0, fx, fy);
                    for (int c = 0; c < C; c++)
                        o[c] = bilerp_(levels[i0[0] * C + c], levels[i0[1] * C + c], levels[i0[2] * C + c], levels[i0[3] * C + c], fx, fy);
                }
            }
    });
}

// Texture backward.  mode 0 ("ref") reproduces texturemap.cu:229-284 including the uvDA
// formula of calculateLevelWithJacob (:160-197, A-2).  mode 1 ("full") writes the exact
// d level / d uvDA chain (zero outside the blended range), for finite-difference agreement.
// Texel grads accumulate (atomics of :249-267), guv/guvDA are stored (:281-282).
// pinned: central differences w.r.t. texels and uv, adjoint identity for texel grads, full-mode
// uvDA grad vs central differences (test_texture_bwd_adjoint_and_fd).
template <class T>
static void tex_bwd(int W, int H, int views, int tw, int th, int C, int L, const T* levels, const T* rast, const T* uv, const T* uvDA,
                    const T* gout, T* glevels, T* guv, T* guvDA, int mode) {
    std::vector<size_t> off(L + 1);
    for (int l = 0; l <= L; l++) off[l] = level_offset(tw, th, C, l);
    Scatter<T> sc(glevels, off[L]);
    parallel_for(views * H, [&](int r0, int r1, int thr) {
        T* gl_ = sc.buf(thr);
        for (int row = r0; row < r1; row++)
            for (int px = 0; px < W; px++) {
                size_t p = (size_t)row * W + px;
                if (rast[4 * p + 3] < T(1)) continue;
                T ux = uv[2 * p], uy = uv[2 * p + 1];
                if (L <= 1) {
                    int i0[4]; T fx, fy;
                    index_fetch<T>(tw, th, 0, ux, uy, i0, fx, fy);
                    T gu = 0, gv = 0;
                    for (int c = 0; c < C; c++) {
                        T g = gout[p * C + c];
                        if (gl_) {
                            gl_[i0[0] * C + c] += (T(1) - fx) * (T(1) - fy) * g;
                            gl_[i0[1] * C + c] += fx * (T(1) - fy) * g;
                            gl_[i0[2] * C + c] += (T(1) - fx) * fy * g;
                            gl_[i0[3] * C + c] += fx * fy * g;
                        }
                        T t00 = levels[i0[0] * C + c], t01 = levels[i0[1] * C + c], t10 = levels[i0[2] * C + c], t11 = levels[i0[3] * C + c];
                        T u = lerp_(t01 - t00, t11 - t10, fy) * (T)tw;
                        T v = lerp_(t10 - t00, t11 - t01, fx) * (T)th;
                        gu += u * g; gv += v * g;
                    }
                    if (guv) { guv[2 * p] = gu; guv[2 * p + 1] = gv; }
                    continue;
                }
                const T* da = uvDA + 4 * p;
                int l0, l1; T f, level;
                calc_level<T>(da, tw, th, L, l0, l1, f, level);
                // d level / d uvDA
                T dl[4];
                {
                    T dsdx = da[0] * tw, dsdy = da[1] * tw, dtdx = da[2] * th, dtdy = da[3] * th;
                    T s2 = dsdx * dsdx + dsdy * dsdy, t2 = dtdx * dtdx + dtdy * dtdy;
                    T a = dsdx * dtdy - dtdx * dsdy;
                    T b = T(0.5f) * (s2 + t2);
                    T pm = dsdx - dtdy, qm = dsdy + dtdx, pp = dsdx + dtdy, qp = dsdy - dtdx;
                    T c2 = T(0.25f) * ((pm * pm + qm * qm) * (pp * pp + qp * qp));
                    T c = std::sqrt(c2);
                    if (mode == 0) {
                        T d = b * c + c2;
                        if (std::fabs((double)d) > 1e-6) {
                            d = T(0.72134752f) / d;
                            T bc = b + c;
                            dl[0] = d * (bc * dsdx - a * dtdy); dl[1] = d * (bc * dsdy + a * dtdx);
                            dl[2] = (bc * dtdx + a * dsdy); dl[3] = (bc * dtdy - a * dsdx);
                        } else if (std::fabs((double)b) > 1e-6) {
                            d = T(1) / b;
                            dl[0] = d * dsdx; dl[1] = d * dsdy; dl[2] = d * dtdx; dl[3] = d * dtdy;
                        } else { dl[0] = dl[1] = dl[2] = dl[3] = 0; }
                    } else {
                        // level = 0.5*log2(b + c); d level = (db + (b db - a da)/c) / (2 ln2 (b + c))
                        T k = T(1) / (T(2) * (T)0.69314718055994530942 * (b + c));
                        bool live = level > 0 && level < (T)(L - 1) && c > T(0) && std::isfinite((double)k);
                        if (live) {
                            T ic = T(1) / c;
                            // partials w.r.t. dsdx, dsdy, dtdx, dtdy
                            T p0 = (dsdx + (b * dsdx - a * dtdy) * ic) * k;
                            T p1 = (dsdy + (b * dsdy + a * dtdx) * ic) * k;
                            T p2 = (dtdx + (b * dtdx + a * dsdy) * ic) * k;
                            T p3 = (dtdy + (b * dtdy - a * dsdx) * ic) * k;
                            dl[0] = p0 * tw; dl[1] = p1 * tw; dl[2] = p2 * th; dl[3] = p3 * th;
                        } else { dl[0] = dl[1] = dl[2] = dl[3] = 0; }
                    }
                }
                int i0[4], i1[4]; T f0x, f0y, f1x, f1y;
                index_fetch<T>(tw, th, l0, ux, uy, i0, f0x, f0y);
                index_fetch<T>(tw, th, l1, ux, uy, i1, f1x, f1y);
                const T* t0 = levels + off[l0]; const T* t1 = levels + off[l1];
                T* g0 = gl_ ? gl_ + off[l0] : nullptr; T* g1 = gl_ ? gl_ + off[l1] : nullptr;
                T gu = 0, gv = 0, gl = 0;
                for (int c = 0; c < C; c++) {
                    T g = gout[p * C + c];
                    if (g0) {
                        g0[i0[0] * C + c] += (T(1) - f) * (T(1) - f0x) * (T(1) - f0y) * g;
                        g0[i0[1] * C + c] += (T(1) - f) * f0x * (T(1) - f0y) * g;
                        g0[i0[2] * C + c] += (T(1) - f) * (T(1) - f0x) * f0y * g;
                        g0[i0[3] * C + c] += (T(1) - f) * f0x * f0y * g;
                    }
                    T t00 = t0[i0[0] * C + c], t01 = t0[i0[1] * C + c], t10 = t0[i0[2] * C + c], t11 = t0[i0[3] * C + c];
                    T u = lerp_(t01 - t00, t11 - t10, f0y) * (T)(tw >> l0);
                    T v = lerp_(t10 - t00, t11 - t01, f0x) * (T)(th >> l0);
                    if (f > 0) {
                        T lv = bilerp_(t00, t01, t10, t11, f0x, f0y);
                        if (g1) {
                            g1[i1[0] * C + c] += f * (T(1) - f1x) * (T(1) - f1y) * g;
                            g1[i1[1] * C + c] += f * f1x * (T(1) - f1y) * g;
                            g1[i1[2] * C + c] += f * (T(1) - f1x) * f1y * g;
                            g1[i1[3] * C + c] += f * f1x * f1y * g;
                        }
                        T s00 = t1[i1[0] * C + c], s01 = t1[i1[1] * C + c], s10 = t1[i1[2] * C + c], s11 = t1[i1[3] * C + c];
                        u = lerp_(u, lerp_(s01 - s00, s11 - s10, f1y) * (T)(tw >> l1), f);
                        v = lerp_(v, lerp_(s10 - s00, s11 - s01, f1x) * (T)(th >> l1), f);
                        gl += (bilerp_(s00, s01, s10, s11, f1x, f1y) - lv) * g;
                    }
                    gu += u * g; gv += v * g;
                }
                if (guv) { guv[2 * p] = gu; guv[2 * p + 1] = gv; }
                if (guvDA) for (int k = 0; k < 4; k++) guvDA[4 * p + k] = gl * dl[k];
            }
    });
    sc.merge();
}

// buildMIP/downSampling (buffer.cu:449-484): 2x2 box mean, levels 1..L-1.
// pinned: constant invariance and closed-form block means (test_mip_build_and_fold_adjoint).
template <class T>
static void mip_build(int tw, int th, int C, int L, T* levels) {
    for (int l = 1; l < L; l++) {
        const T* src = levels + level_offset(tw, th, C, l - 1);
        T* dst = levels + level_offset(tw, th, C, l);
        int w = tw >> l, h = th >> l, sw = tw >> (l - 1);
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++)
                for (int c = 0; c < C; c++) {
                    T p00 = src[((size_t)(2 * y) * sw + 2 * x) * C + c], p01 = src[((size_t)(2 * y) * sw + 2 * x + 1) * C + c];
                    T p10 = src[((size_t)(2 * y + 1) * sw + 2 * x) * C + c], p11 = src[((size_t)(2 * y + 1) * sw + 2 * x + 1) * C + c];
                    dst[((size_t)y * w + x) * C + c] = (p00 + p01 + p10 + p11) * T(0.25f);
                }
    }
}

// gradSumup/gardAddThrough (buffer.cu:615-648): coarse-to-fine, each coarse grad added to its
// four children with AddNaNcheck.  quarter=1 scales by 1/4 (exact adjoint of mip_build, A-3);
// quarter=0 is the reference's x1.
// pinned: adjoint identity <build(x), g> = <x, fold(g)> (test_mip_build_and_fold_adjoint).
template <class T>
static void mip_fold(int tw, int th, int C, int L, T* glevels, int quarter) {
    for (int l = L - 1; l > 0; l--) {
        const T* src = glevels + level_offset(tw, th, C, l);
        T* dst = glevels + level_offset(tw, th, C, l - 1);
        int w = tw >> l, h = th >> l, dw = tw >> (l - 1);
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++)
                for (int c = 0; c < C; c++) {
                    T g = src[((size_t)y * w + x) * C + c];
                    if (quarter) g = g * T(0.25f);
                    add_nan_check(dst[((size_t)(2 * y) * dw + 2 * x) * C + c], g);
                    add_nan_check(dst[((size_t)(2 * y) * dw + 2 * x + 1) * C + c], g);
                    add_nan_check(dst[((size_t)(2 * y + 1) * dw + 2 * x) * C + c], g);
                    add_nan_check(dst[((size_t)(2 * y + 1) * dw + 2 * x + 1) * C + c], g);
                }
    }
}

// ------------------------------------------------------------------------------------
// NormalAxis (material.cu:25-69): out = rot3x3 * normalize(n0' u + n1' v + n2), the effective
// semantics of the reference (A-8).  rot column-major 4x4 (16 values).
// pinned: closed form (test_normal_and_view_axis_closed_form).
template <class T>
static void normal_axis(int W, int H, int views, const T* rast, const T* normals, const u32* nidx, const T* rot, T* out) {
    parallel_for(views * H, [&](int r0, int r1, int) {
        for (int row = r0; row < r1; row++)
            for (int px = 0; px < W; px++) {
                size_t p = (size_t)row * W + px;
                const T* r = rast + 4 * p; T* o = out + 3 * p;
                int id = (int)r[3] - 1;
                if (id < 0) { o[0] = o[1] = o[2] = 0; continue; }
                const T* n2 = normals + 3 * (size_t)nidx[3 * id + 2];
                const T* n0 = normals + 3 * (size_t)nidx[3 * id];
                const T* n1 = normals + 3 * (size_t)nidx[3 * id + 1];
                T N[3];
                for (int k = 0; k < 3; k++) N[k] = (n0[k] - n2[k]) * r[0] + (n1[k] - n2[k]) * r[1] + n2[k];
                T d = T(1) / std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
                N[0] *= d; N[1] *= d; N[2] *= d;
                o[0] = rot[0] * N[0] + rot[4] * N[1] + rot[8] * N[2];
                o[1] = rot[1] * N[0] + rot[5] * N[1] + rot[9] * N[2];
                o[2] = rot[2] * N[0] + rot[6] * N[1] + rot[10] * N[2];
            }
    });
}

// ViewAxis (material.cu:92-118): out = normalize(pvinv[0]*sx + pvinv[1]*sy + pvinv[2]) with
// sx = px/W*2-1 (no half-pixel offset, A-9); pvinv column-major 3x3, one per view.  Written
// for every pixel.
// pinned: closed form (test_normal_and_view_axis_closed_form).
template <class T>
static void view_axis(int W, int H, int views, const T* pvinv, T* out) {
    parallel_for(views * H, [&](int r0, int r1, int) {
        for (int row = r0; row < r1; row++) {
            int s = row / H, py = row % H;
            const T* m = pvinv + 9 * s;
            for (int px = 0; px < W; px++) {
                size_t p = (size_t)row * W + px;
                T sx = (T)px / (T)W * T(2) - T(1), sy = (T)py / (T)H * T(2) - T(1);
                T v[3];
                for (int k = 0; k < 3; k++) v[k] = m[k] * sx + m[3 + k] * sy + m[6 + k];
                T d = T(1) / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                out[3 * p] = v[0] * d; out[3 * p + 1] = v[1] * d; out[3 * p + 2] = v[2] * d;
            }
        }
    });
}

// SGSpecular forward (material.cu:145-183) and backward (:203-249, roughness only).
// pinned: forward by textbook optics -- Householder reflection axis, the Fresnel equations and
// GGX Smith G1 (test_sg_specular_forward_closed_forms); roughness grad vs central differences
// (test_sg_specular_roughness_fd).
template <class T>
static void sg_spec_terms(const T* n, const T* view, T m, T ior, T* p, T& amp, T& sharp,
                          T* dAmp_dm, T* dSharp_dm) {
    T v[3] = {-view[0], -view[1], -view[2]};
    T vdn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
    for (int k = 0; k < 3; k++) p[k] = T(2) * vdn * n[k] - v[k];
    T m2 = m * m;
    T vn = std::max(vdn, T(0));
    T vn2 = vn * vn;
    T iGv = std::sqrt(vn2 + m2 * (T(1) - vn2));
    T Gv = (T)(1.0 / dmax((double)(vn + iGv), 1e-6));
    T pdn = p[0] * n[0] + p[1] * n[1] + p[2] * n[2];
    T pn = std::max(pdn, T(0));
    T pn2 = pn * pn;
    T iGp = std::sqrt(pn2 + m2 * (T(1) - pn2));
    T Gp = (T)(1.0 / dmax((double)(pn + iGp), 1e-6));
    T g = std::sqrt(ior * ior - T(1) + pn2);
    T gp = g + pn, gm = g - pn;
    T f0 = gm / gp;
    T f1 = (pn * gp - T(1)) / (pn * gm + T(1));
    T F = f0 * f0 * (T(1) + f1 * f1);
    amp = (T)(F * Gp * Gv * pn / dmax((double)m2, 1e-6));
    sharp = (T)(0.5f / dmax((double)(m2 * vn), 1e-6));
    if (dAmp_dm) {
        T dGvdm = (T)(-m * (T(1) - vn2) / dmax((double)iGv, 1e-6) * Gv * Gv);
        T dGpdm = (T)(-m * (T(1) - pn2) / dmax((double)iGp, 1e-6) * Gp * Gp);
        *dAmp_dm = (T)(F * pn * (-2.0 * Gp * Gv / dmax((double)(m2 * m), 1e-6) + (Gp * dGvdm + Gv * dGpdm) / dmax((double)m2, 1e-6)));
        *dSharp_dm = (T)(-1.0 / dmax((double)(m2 * m * vn), 1e-6));
    }
}

template <class T>
static void sg_spec_fwd(int W, int H, int views, const T* rast, const T* normal, const T* view, const T* rough, T ior,
                        T* outAxis, T* outSharp, T* outAmp) {
    parallel_for(views * H, [&](int r0, int r1, int) {
        for (int row = r0; row < r1; row++)
            for (int px = 0; px < W; px++) {
                size_t p = (size_t)row * W + px;
                if ((int)rast[4 * p + 3] - 1 < 0) { outAxis[3 * p] = outAxis[3 * p + 1] = outAxis[3 * p + 2] = 0; outSharp[p] = 0; outAmp[p] = 0; continue; }
                T amp, sh;
                sg_spec_terms<T>(normal + 3 * p, view + 3 * p, rough[p], ior, outAxis + 3 * p, amp, sh, nullptr, nullptr);
                outAmp[p] = amp; outSharp[p] = sh;
            }
    });
}

// Roughness grad: AddNaNcheck-accumulated as in material.cu:247-248.  The reference's
// dAmplitude/dm is used as is (its -2 Gp Gv / m^3 term is d(1/m^2)/dm = -2/m^3).
template <class T>
static void sg_spec_bwd(int W, int H, int views, const T* rast, const T* normal, const T* view, const T* rough, T ior,
                        const T* gAmp, const T* gSharp, T* grough) {
    parallel_for(views * H, [&](int r0, int r1, int) {
        for (int row = r0; row < r1; row++)
            for (int px = 0; px < W; px++) {
                size_t p = (size_t)row * W + px;
                if ((int)rast[4 * p + 3] - 1 < 0) continue;
                T ax[3], amp, sh, dA, dS;
                sg_spec_terms<T>(normal + 3 * p, view + 3 * p, rough[p], ior, ax, amp, sh, &dA, &dS);
                add_nan_check(grough[p], gAmp[p] * dA);
                add_nan_check(grough[p], gSharp[p] * dS);
            }
    });
}

// SphericalGaussian forward (material.cu:285-323).
//   g(l, c) = (exp(l - c) - exp(-l - c)) / max(l, 1e-6)
//   diffenv[k] += g(|2.133 n + l_i mu_i|, 2.133 + l_i) * 2.34 * a_i[k]
//   specenv[k] += g(|l_s p + l_i mu_i|, l_s + l_i) * A_s * a_i[k]     (AddNaNcheck, lobe order)
//   out[k] = diffuse[k] * diffenv[k] + specenv[k]
// pinned: aligned-axis closed form g = (1 - e^{-2l})/l (test_sg_aligned_closed_form).
static const float SG_DIFF_SHARP = 2.133f, SG_DIFF_AMP = 2.34f;

template <class T>
static void sg_fwd(int W, int H, int views, int C, int K, const T* rast, const T* axis, const T* sharp, const T* amp,
                   const T* normal, const T* diffuse, const T* sAxis, const T* sSharp, const T* sAmp,
                   T* out, T* diffenv, T* specenv) {
    parallel_for(views * H, [&](int r0, int r1, int) {
        for (int row = r0; row < r1; row++)
            for (int px = 0; px < W; px++) {
                size_t p = (size_t)row * W + px;
                T* o = out + p * C; T* de = diffenv + p * C; T* se = specenv + p * C;
                for (int k = 0; k < C; k++) { o[k] = 0; de[k] = 0; se[k] = 0; }
                if ((int)rast[4 * p + 3] - 1 < 0) continue;
                const T* n = normal + 3 * p;
                T ds = (T)SG_DIFF_SHARP, da = (T)SG_DIFF_AMP;
                T dn[3] = {n[0] * ds, n[1] * ds, n[2] * ds};
                const T* pp = sAxis + 3 * p;
                T ss = sSharp[p], sa = sAmp[p];
                T sn[3] = {pp[0] * ss, pp[1] * ss, pp[2] * ss};
                for (int i = 0; i < K; i++) {
                    T lam = sharp[i];
                    T g[3] = {axis[3 * i] * lam, axis[3 * i + 1] * lam, axis[3 * i + 2] * lam};
                    T w0 = dn[0] + g[0], w1 = dn[1] + g[1], w2 = dn[2] + g[2];
                    T dl = std::sqrt(w0 * w0 + w1 * w1 + w2 * w2);
                    T dg = (T)((std::exp(dl - ds - lam) - std::exp(-dl - ds - lam)) / dmax((double)dl, 1e-6) * da);
                    T v0 = sn[0] + g[0], v1 = sn[1] + g[1], v2 = sn[2] + g[2];
                    T sl = std::sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                    T sg = (T)((std::exp(sl - ss - lam) - std::exp(-sl - ss - lam)) / dmax((double)sl, 1e-6) * sa);
                    for (int k = 0; k < C; k++) {
                        T a = amp[i * C + k];
                        add_nan_check(de[k], dg * a);
                        add_nan_check(se[k], sg * a);
                    }
                }
                for (int k = 0; k < C; k++) o[k] = diffuse[p * C + k] * de[k] + se[k];
            }
    });
}

// SphericalGaussian backward, race-free and with the exact derivatives (A-7: the reference
// accumulates lobe grads non-atomically, uses lambda*mu.n in d l / d lambda, and indexes the
// per-pixel spec grads by lobe/channel).  Lobe grads are summed per thread and merged in
// thread order.  Per pixel (w = l_x x + l_i mu_i, l = |w|, e1 = e^{l-c}, e2 = e^{-l-c}):
//   dg/dl = ((1 - 1/l) e1 + (1 + 1/l) e2) / l,  dg/dc = -g
// gdiffuse[k] += gout[k] * diffenv[k] (accumulated), grad.normal/spec grads per pixel.
// pinned: central differences for every parameter (test_sg_bwd_finite_differences); grad.diffuse equals the
// reference's race-free line material.cu:429.
template <class T>
static void sg_bwd(int W, int H, int views, int C, int K, const T* rast, const T* axis, const T* sharp, const T* amp,
                   const T* normal, const T* diffuse, const T* sAxis, const T* sSharp, const T* sAmp, const T* gout,
                   T* gdiffuse, T* gaxis, T* gsharp, T* gamp, T* gnormal, T* gsAxis, T* gsSharp, T* gsAmp) {
    Scatter<T> scA(gaxis, (size_t)K * 3), scS(gsharp, (size_t)K), scM(gamp, (size_t)K * C);
    parallel_for(views * H, [&](int r0, int r1, int th) {
        T* ga = scA.buf(th); T* gs = scS.buf(th); T* gm = scM.buf(th);
        std::vector<T> Sd(K), Ss(K);
        for (int row = r0; row < r1; row++)
            for (int px = 0; px < W; px++) {
                size_t p = (size_t)row * W + px;
                if ((int)rast[4 * p + 3] - 1 < 0) continue;
                const T* n = normal + 3 * p;
                T ds = (T)SG_DIFF_SHARP, da = (T)SG_DIFF_AMP;
                T dn[3] = {n[0] * ds, n[1] * ds, n[2] * ds};
                const T* pp = sAxis + 3 * p;
                T ss = sSharp[p], sa = sAmp[p];
                T sn[3] = {pp[0] * ss, pp[1] * ss, pp[2] * ss};
                T gn[3] = {0, 0, 0}, gp[3] = {0, 0, 0}, gss = 0, gsa = 0;
                for (int i = 0; i < K; i++) {
                    T lam = sharp[i];
                    const T* mu = axis + 3 * i;
                    T g[3] = {mu[0] * lam, mu[1] * lam, mu[2] * lam};
                    // diffuse lobe
                    T w[3] = {dn[0] + g[0], dn[1] + g[1], dn[2] + g[2]};
                    T l = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
                    T il = (T)(1.0 / dmax((double)l, 1e-6));
                    T e1 = std::exp(l - ds - lam), e2 = std::exp(-l - ds - lam);
                    T dgv = (e1 - e2) * il;                     // g
                    T dgdl = ((T(1) - il) * e1 + (T(1) + il) * e2) * il;
                    // specular lobe
                    T q[3] = {sn[0] + g[0], sn[1] + g[1], sn[2] + g[2]};
                    T sl = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
                    T isl = (T)(1.0 / dmax((double)sl, 1e-6));
                    T f1 = std::exp(sl - ss - lam), f2 = std::exp(-sl - ss - lam);
                    T sgv = (f1 - f2) * isl;
                    T sgdl = ((T(1) - isl) * f1 + (T(1) + isl) * f2) * isl;
                    T sdiff = 0, ssum = 0;
                    for (int k = 0; k < C; k++) {
                        T go = gout[p * C + k], a = amp[i * C + k];
                        sdiff += go * a * diffuse[p * C + k];
                        ssum += go * a;
                        if (gm) gm[i * C + k] += go * (dgv * da * diffuse[p * C + k] + sgv * sa);
                    }
                    // d/d(l_i mu_i): Sd*da*dgdl*w/l + Ss*sa*sgdl*q/sl
                    T cd = sdiff * da * dgdl * il, cs = ssum * sa * sgdl * isl;
                    T gv3[3] = {cd * w[0] + cs * q[0], cd * w[1] + cs * q[1], cd * w[2] + cs * q[2]};
                    if (ga) for (int k = 0; k < 3; k++) ga[3 * i + k] += gv3[k] * lam;
                    if (gs) {
                        T mdw = mu[0] * w[0] + mu[1] * w[1] + mu[2] * w[2];
                        T mdq = mu[0] * q[0] + mu[1] * q[1] + mu[2] * q[2];
                        gs[i] += sdiff * da * (dgdl * mdw * il - dgv) + ssum * sa * (sgdl * mdq * isl - sgv);
                    }
                    for (int k = 0; k < 3; k++) { gn[k] += cd * w[k] * ds; gp[k] += cs * q[k] * ss; }
                    T pdq = pp[0] * q[0] + pp[1] * q[1] + pp[2] * q[2];
                    gss += ssum * sa * (sgdl * pdq * isl - sgv);
                    gsa += ssum * sgv;
                }
                if (gdiffuse) {
                    // diffenv recomputed in lobe order as the forward does
                    for (int k = 0; k < C; k++) {
                        T de = 0;
                        for (int i = 0; i < K; i++) {
                            T lam = sharp[i];
                            T w0 = dn[0] + axis[3 * i] * lam, w1 = dn[1] + axis[3 * i + 1] * lam, w2 = dn[2] + axis[3 * i + 2] * lam;
                            T dl = std::sqrt(w0 * w0 + w1 * w1 + w2 * w2);
                            T dg = (T)((std::exp(dl - ds - lam) - std::exp(-dl - ds - lam)) / dmax((double)dl, 1e-6) * da);
                            add_nan_check(de, dg * amp[i * C + k]);
                        }
                        add_nan_check(gdiffuse[p * C + k], gout[p * C + k] * de);
                    }
                }
                if (gnormal) for (int k = 0; k < 3; k++) gnormal[3 * p + k] += gn[k];
                if (gsAxis) for (int k = 0; k < 3; k++) gsAxis[3 * p + k] += gp[k];
                if (gsSharp) gsSharp[p] += gss;
                if (gsAmp) gsAmp[p] += gsa;
            }
    });
    scA.merge(); scS.merge(); scM.merge();
}

// ------------------------------------------------------------------------------------
// Antialias (antialias.cu:17-95 forward, :196-300 backward), race-free gather form.
//
// A pair (q, n) exists for each pixel q with a left (n = q-1) or lower (n = q-W) neighbour
// holding a different triangle ID.  "p" is the covered pixel, or the farther one when both
// are covered (A-6); "o" is the other one.  For each edge (va,vb) of p's triangle, in pixel
// space relative to p's centre: if the edge crosses segment p->o and is more perpendicular
// than parallel to it, alpha = a/(D +- 1e-3) - 1/2 and alpha*(in[p]-in[o]) is added to o
// (alpha > 0) or p (antialias.cu:18-36).  out = in + all contributions (A-4 fixed); the sum
// for a pixel is taken over its pairs in the order: left, lower, right, upper.
// The backward mirrors the forward's p/o selection (A-5 fixed) and uses the exact derivative
// of alpha, including the w term (the reference's w grad uses the pixel-space offset instead
// of x/w, antialias.cu:237-238).
// pinned: single straight edge closed form (test_aa_straight_edge_closed_form), mirror symmetry
// (test_aa_mirror_symmetry), central differences w.r.t. clip coords and input colours
// (test_aa_bwd_finite_differences).
struct AAEdge { int valid; double alpha; int target_is_o; int va, vb; double ia; double pax, pay, pbx, pby; int horiz; float sg; };

template <class T>
struct AAPair {
    int p, o; int tri; int nE; AAEdge e[3];
};

template <class T>
static bool aa_pair(int W, int H, const T* clipv, const u32* idx, const T* rastv, int px, int py, int vertical, AAPair<T>& out) {
    size_t pid = (size_t)py * W + px;
    size_t nid;
    if (!vertical) { if (px <= 0) return false; nid = pid - 1; } else { if (py <= 0) return false; nid = pid - W; }
    T tz = rastv[4 * pid + 2], tid = rastv[4 * pid + 3];
    T nz = rastv[4 * nid + 2], nidv = rastv[4 * nid + 3];
    if (nidv == tid) return false;
    size_t pp = (int)tid ? pid : nid;
    size_t op = (int)nidv ? pid : nid;
    if (pp == op) { if (tz > nz) op = nid; else pp = nid; }
    float fx, fy, ox, oy;
    if (!vertical) { fx = (float)px + (pp > op ? .5f : -.5f); fy = (float)py + .5f; ox = pp > op ? -1.f : 1.f; oy = 0.f; }
    else { fx = (float)px + .5f; fy = (float)py + (pp > op ? .5f : -.5f); ox = 0.f; oy = pp > op ? -1.f : 1.f; }
    out.p = (int)pp; out.o = (int)op;
    int id = (int)rastv[4 * pp + 3] - 1;
    out.tri = id; out.nE = 0;
    T sx = (T)(W / 2.f), sy = (T)(H / 2.f);
    u32 vi[3] = {idx[3 * id], idx[3 * id + 1], idx[3 * id + 2]};
    T vx[3], vy[3];
    for (int k = 0; k < 3; k++) {
        const T* c = clipv + 4 * (size_t)vi[k];
        T iw = T(1) / c[3];
        vx[k] = (c[0] * iw + T(1)) * sx - (T)fx;
        vy[k] = (c[1] * iw + T(1)) * sy - (T)fy;
    }
    for (int k = 0; k < 3; k++) {
        int a_ = k, b_ = (k + 1) % 3;
        AAEdge& E = out.e[k];
        E.valid = 0; E.va = (int)vi[a_]; E.vb = (int)vi[b_];
        T pax = vx[a_], pay = vy[a_], pbx = vx[b_], pby = vy[b_];
        T a = pax * pby - pay * pbx;
        T oa = (pax - ox) * (pby - oy) - (pay - oy) * (pbx - ox);
        if (a * oa > T(0)) continue;
        T ex = pax - pbx, ey = pay - pby;
        T n = (ox + oy) * (ox - oy);
        if ((ex + ey) * (ex - ey) * n > T(0)) continue;
        T sg = ox + oy;
        ex *= sg; ey *= sg;
        T D = n > T(0) ? -ey : ex;
        T ia = (T)(1.0 / ((double)D + (D > 0 ? 1e-3 : -1e-3)));
        T alpha = a * ia - T(0.5f);
        E.valid = 1; E.alpha = (double)alpha; E.target_is_o = alpha > 0 ? 1 : 0;
        E.ia = (double)ia; E.pax = (double)pax; E.pay = (double)pay; E.pbx = (double)pbx; E.pby = (double)pby;
        E.horiz = n > T(0) ? 1 : 0; E.sg = (float)sg;
    }
    return true;
}

template <class T>
static void aa_fwd(int W, int H, int views, int C, int V, const T* clip, const u32* idx, const T* rast, const T* in, T* out) {
    parallel_for(views * H, [&](int r0, int r1, int) {
        for (int row = r0; row < r1; row++) {
            int s = row / H, py = row % H;
            const T* clipv = clip + (size_t)s * V * 4;
            const T* rastv = rast + (size_t)s * W * H * 4;
            const T* inv = in + (size_t)s * W * H * C;
            for (int px = 0; px < W; px++) {
                size_t q = (size_t)py * W + px;
                T acc[16];
                for (int c = 0; c < C; c++) acc[c] = inv[q * C + c];
                // pairs touching q: (q,left) (q,down) (right,q) (up,q)
                int cand[4][3] = {{px, py, 0}, {px, py, 1}, {px + 1, py, 0}, {px, py + 1, 1}};
                for (int k = 0; k < 4; k++) {
                    int cx = cand[k][0], cy = cand[k][1];
                    if (cx >= W || cy >= H) continue;
                    AAPair<T> P;
                    if (!aa_pair<T>(W, H, clipv, idx, rastv, cx, cy, cand[k][2], P)) continue;
                    for (int e = 0; e < 3; e++) {
                        const AAEdge& E = P.e[e];
                        if (!E.valid) continue;
                        int tgt = E.target_is_o ? P.o : P.p;
                        if ((size_t)tgt != q) continue;
                        T al = (T)E.alpha;
                        for (int c = 0; c < C; c++) acc[c] += al * (inv[(size_t)P.p * C + c] - inv[(size_t)P.o * C + c]);
                    }
                }
                for (int c = 0; c < C; c++) out[((size_t)s * W * H + q) * C + c] = acc[c];
            }
        }
    });
}

template <class T>
static void aa_bwd(int W, int H, int views, int C, int V, const T* clip, const u32* idx, const T* rast, const T* in, const T* gout,
                   T* gin, T* gclip) {
    Scatter<T> sc(gclip, (size_t)views * V * 4);
    parallel_for(views * H, [&](int r0, int r1, int th) {
        T* gcl = sc.buf(th);
        for (int row = r0; row < r1; row++) {
            int s = row / H, py = row % H;
            const T* clipv = clip + (size_t)s * V * 4;
            const T* rastv = rast + (size_t)s * W * H * 4;
            const T* inv = in + (size_t)s * W * H * C;
            const T* gov = gout + (size_t)s * W * H * C;
            T* gcs = gcl ? gcl + (size_t)s * V * 4 : nullptr;
            T sx = (T)(W / 2.f), sy = (T)(H / 2.f);
            for (int px = 0; px < W; px++) {
                size_t q = (size_t)py * W + px;
                T acc[16];
                for (int c = 0; c < C; c++) acc[c] = gov[q * C + c];
                int cand[4][3] = {{px, py, 0}, {px, py, 1}, {px + 1, py, 0}, {px, py + 1, 1}};
                for (int k = 0; k < 4; k++) {
                    int cx = cand[k][0], cy = cand[k][1];
                    if (cx >= W || cy >= H) continue;
                    AAPair<T> P;
                    if (!aa_pair<T>(W, H, clipv, idx, rastv, cx, cy, cand[k][2], P)) continue;
                    for (int e = 0; e < 3; e++) {
                        const AAEdge& E = P.e[e];
                        if (!E.valid) continue;
                        int tgt = E.target_is_o ? P.o : P.p;
                        T al = (T)E.alpha;
                        if ((size_t)P.p == q) for (int c = 0; c < C; c++) acc[c] += al * gov[(size_t)tgt * C + c];
                        if ((size_t)P.o == q) for (int c = 0; c < C; c++) acc[c] -= al * gov[(size_t)tgt * C + c];
                        // clip-coordinate grads: the pair is owned by its "current" pixel (cx,cy)
                        if (gcs && k < 2) {
                            T d = 0;
                            for (int c = 0; c < C; c++) d += gov[(size_t)tgt * C + c] * (inv[(size_t)P.p * C + c] - inv[(size_t)P.o * C + c]);
                            T ia = (T)E.ia, a = (T)(E.pax * E.pby - E.pay * E.pbx);
                            T pax = (T)E.pax, pay = (T)E.pay, pbx = (T)E.pbx, pby = (T)E.pby, sg = (T)E.sg;
                            // dalpha/dpa, dalpha/dpb = ia*da - a*ia^2*dD
                            T dDax, dDay, dDbx, dDby;
                            if (E.horiz) { dDax = 0; dDay = -sg; dDbx = 0; dDby = sg; }
                            else { dDax = sg; dDay = 0; dDbx = -sg; dDby = 0; }
                            T k2 = a * ia * ia;
                            T gax = d * (ia * pby - k2 * dDax), gay = d * (-ia * pbx - k2 * dDay);
                            T gbx = d * (-ia * pay - k2 * dDbx), gby = d * (ia * pax - k2 * dDby);
                            const int vv[2] = {E.va, E.vb};
                            const T gx[2] = {gax, gbx}, gy[2] = {gay, gby};
                            for (int j = 0; j < 2; j++) {
                                const T* c = clipv + 4 * (size_t)vv[j];
                                T iw = T(1) / c[3];
                                T gvx = gx[j] * sx * iw, gvy = gy[j] * sy * iw;
                                T* g = gcs + 4 * (size_t)vv[j];
                                g[0] += gvx; g[1] += gvy; g[3] += -(gvx * c[0] + gvy * c[1]) * iw;
                            }
                        }
                    }
                }
                for (int c = 0; c < C; c++) gin[((size_t)s * W * H + q) * C + c] = acc[c];
            }
        }
    });
    sc.merge();
}

// ------------------------------------------------------------------------------------
// MSE (loss.cu:22-57): grad = pred - target, loss = 0.5 * sum(grad^2) (summed in double here;
// the reference's strided tree is exact only for power-of-two sizes, A-13).
template <class T>
static double mse(size_t n, const T* pred, const T* target, T* grad) {
    double s = 0;
    for (size_t i = 0; i < n; i++) { T g = pred[i] - target[i]; grad[i] = g; s += (double)g * (double)g; }
    return 0.5 * s;
}

// Adam (optimize.cu:95-115): host bumps rhomt/rhovt before the launch, so step t uses
// beta^(t+1) (A-10); m, v updated in double and stored as T; update via AddNaNcheck.
// pinned: brute-force FP64 recurrence (test_adam_recurrence_fp64).
template <class T>
static void adam(size_t n, T* param, const T* grad, T* m, T* v, double rhom, double rhov, double rhomt, double rhovt, double eta, double eps) {
    for (size_t i = 0; i < n; i++) {
        m[i] = (T)(rhom * (double)m[i] + (1 - rhom) * (double)grad[i]);
        v[i] = (T)(rhov * (double)v[i] + (1 - rhov) * (double)grad[i] * (double)grad[i]);
        double mh = (double)m[i] / (1 - rhomt);
        double vh = (double)v[i] / (1 - rhovt);
        add_nan_check(param[i], (T)(-mh * eta / (std::sqrt(vh) + eps)));
    }
}

// clampKernel over level 0 (buffer.cu:65-72, Texture::clamp buffer.cu:588-594).
template <class T>
static void clamp_(size_t n, T* x, T lo, T hi) { for (size_t i = 0; i < n; i++) x[i] = x[i] < lo ? lo : x[i] > hi ? hi : x[i]; }

// ------------------------------------------------------------------------------------
// extern "C" wrappers
#define ORC_BOTH(MACRO) MACRO(float, f32) MACRO(double, f64)

#define ORC_WRAP(T, S) \
extern "C" void orc_project_fwd_##S(int V, int views, const T* pos, const T* mats, T* clip) { project_fwd<T>(V, views, pos, mats, clip); } \
extern "C" void orc_project_bwd_##S(int V, int views, const T* mats, const T* gclip, T* gpos) { project_bwd<T>(V, views, mats, gclip, gpos); } \
extern "C" void orc_raster_fwd_##S(int W, int H, int views, int V, const T* clip, int T_, const u32* idx, T* rast, T* db) { raster_fwd<T>(W, H, views, V, clip, T_, idx, rast, db); } \
extern "C" void orc_raster_bwd_##S(int W, int H, int views, int V, const T* clip, const u32* idx, const T* rast, const T* grast, T* gclip) { raster_bwd<T>(W, H, views, V, clip, idx, rast, grast, gclip); } \
extern "C" void orc_interp_fwd_##S(int W, int H, int views, int D, const T* rast, const T* db, const T* attr, const u32* idx, T* out, T* outDA) { interp_fwd<T>(W, H, views, D, rast, db, attr, idx, out, outDA); } \
extern "C" void orc_interp_bwd_##S(int W, int H, int views, int D, int V, const T* rast, const T* attr, const u32* idx, const T* gout, T* gattr, T* grast) { interp_bwd<T>(W, H, views, D, rast, attr, idx, gout, gattr, grast, V); } \
extern "C" void orc_tex_fwd_##S(int W, int H, int views, int tw, int th, int C, int L, const T* levels, const T* rast, const T* uv, const T* uvDA, T* out) { tex_fwd<T>(W, H, views, tw, th, C, L, levels, rast, uv, uvDA, out); } \
extern "C" void orc_tex_bwd_##S(int W, int H, int views, int tw, int th, int C, int L, const T* levels, const T* rast, const T* uv, const T* uvDA, const T* gout, T* glevels, T* guv, T* guvDA, int mode) { tex_bwd<T>(W, H, views, tw, th, C, L, levels, rast, uv, uvDA, gout, glevels, guv, guvDA, mode); } \
extern "C" void orc_mip_build_##S(int tw, int th, int C, int L, T* levels) { mip_build<T>(tw, th, C, L, levels); } \
extern "C" void orc_mip_fold_##S(int tw, int th, int C, int L, T* glevels, int quarter) { mip_fold<T>(tw, th, C, L, glevels, quarter); } \
extern "C" void orc_normal_axis_##S(int W, int H, int views, const T* rast, const T* normals, const u32* nidx, const T* rot, T* out) { normal_axis<T>(W, H, views, rast, normals, nidx, rot, out); } \
extern "C" void orc_view_axis_##S(int W, int H, int views, const T* pvinv, T* out) { view_axis<T>(W, H, views, pvinv, out); } \
extern "C" void orc_sg_spec_fwd_##S(int W, int H, int views, const T* rast, const T* normal, const T* view, const T* rough, T ior, T* outAxis, T* outSharp, T* outAmp) { sg_spec_fwd<T>(W, H, views, rast, normal, view, rough, ior, outAxis, outSharp, outAmp); } \
extern "C" void orc_sg_spec_bwd_##S(int W, int H, int views, const T* rast, const T* normal, const T* view, const T* rough, T ior, const T* gAmp, const T* gSharp, T* grough) { sg_spec_bwd<T>(W, H, views, rast, normal, view, rough, ior, gAmp, gSharp, grough); } \
extern "C" void orc_sg_fwd_##S(int W, int H, int views, int C, int K, const T* rast, const T* axis, const T* sharp, const T* amp, const T* normal, const T* diffuse, const T* sAxis, const T* sSharp, const T* sAmp, T* out, T* diffenv, T* specenv) { sg_fwd<T>(W, H, views, C, K, rast, axis, sharp, amp, normal, diffuse, sAxis, sSharp, sAmp, out, diffenv, specenv); } \
extern "C" void orc_sg_bwd_##S(int W, int H, int views, int C, int K, const T* rast, const T* axis, const T* sharp, const T* amp, const T* normal, const T* diffuse, const T* sAxis, const T* sSharp, const T* sAmp, const T* gout, T* gdiffuse, T* gaxis, T* gsharp, T* gamp, T* gnormal, T* gsAxis, T* gsSharp, T* gsAmp) { sg_bwd<T>(W, H, views, C, K, rast, axis, sharp, amp, normal, diffuse, sAxis, sSharp, sAmp, gout, gdiffuse, gaxis, gsharp, gamp, gnormal, gsAxis, gsSharp, gsAmp); } \
extern "C" void orc_aa_fwd_##S(int W, int H, int views, int C, int V, const T* clip, const u32* idx, const T* rast, const T* in, T* out) { aa_fwd<T>(W, H, views, C, V, clip, idx, rast, in, out); } \
extern "C" void orc_aa_bwd_##S(int W, int H, int views, int C, int V, const T* clip, const u32* idx, const T* rast, const T* in, const T* gout, T* gin, T* gclip) { aa_bwd<T>(W, H, views, C, V, clip, idx, rast, in, gout, gin, gclip); } \
extern "C" double orc_mse_##S(size_t n, const T* pred, const T* target, T* grad) { return mse<T>(n, pred, target, grad); } \
extern "C" void orc_adam_##S(size_t n, T* param, const T* grad, T* m, T* v, double rhom, double rhov, double rhomt, double rhovt, double eta, double eps) { adam<T>(n, param, grad, m, v, rhom, rhov, rhomt, rhovt, eta, eps); } \
extern "C" void orc_clamp_##S(size_t n, T* x, T lo, T hi) { clamp_<T>(n, x, lo, hi); }

ORC_BOTH(ORC_WRAP)

// ------------------------------------------------------------------------------------
// One fwd+bwd+Adam fit iteration, in the call order of preset_primitives.cpp:167-207 for the
// nvdiffrast-style chain: project -> rasterize(+DB) -> interpolate(uv,+DA) -> NormalAxis ->
// ViewAxis -> mip texture -> SGSpecular -> SphericalGaussian -> antialias -> MSE, then the
// backward chain antialias -> SG -> texture -> interpolate -> rasterize -> project, then
// gradSumup -> Adam(texture L0) -> clear -> clamp+buildMIP, and Adam on positions and lobes.
// `views` cameras share the mesh, texture and lobes: per-view arrays (mat, pvinv, rough,
// target, out_image) hold `views` consecutive slices, the loss and every parameter gradient
// sum over the views.  orc_fit_step_f32 is the CPU baseline that bench.py times;
// orc_fit_grads_f64 evaluates the same forward and backward in double (the error yardstick).
struct OrcScene {
    int W, H, V, T, C, K, tw, th, L;
    float* pos; const float* normals; const float* uv; const u32* idx;
    const float* mat; const float* rot; const float* pvinv;
    float* tex;            // all levels
    float* axis; float* sharp; float* amp;
    const float* rough;    // views x P roughness (constant map)
    float ior;
    const float* target;   // views x P x C
    int fold_quarter;      // 1: exact adjoint fold
    // Adam state
    float *m_pos, *v_pos, *m_tex, *v_tex, *m_axis, *v_axis, *m_sharp, *v_sharp, *m_amp, *v_amp;
    double lr, b1, b2, eps, b1t, b2t;
    // outputs
    float* out_image; double loss;
    int views;             // 0 or 1: one view
};

template <class T>
struct FitWork {
    std::vector<T> clip, rast, db, uvp, uvDA, nrm, view, diff, sAxis, sSharp, sAmp, color, denv, senv, img, gimg, gcolor,
        gdiff, gtex, guv, guvDA, grast, gclip, gpos, gaxis, gsharp, gamp;
    // double copies of the float inputs (orc_fit_grads_f64 only)
    std::vector<T> in_pos, in_normals, in_uv, in_mat, in_rot, in_pvinv, in_tex, in_axis, in_sharp, in_amp, in_rough, in_target;
};
typedef FitWork<float> OrcWork;

// Forward and backward of one iteration; leaves every parameter gradient in the work buffers
// (texture grads per level, not folded).  Returns the loss.
// cov_clip: see raster_fwd (NULL in the float step).
template <class T>
static double fit_fwd_bwd(const OrcScene* s, FitWork<T>& w, const T* pos, const T* normals, const T* uv, const T* mat, const T* rot,
                          const T* pvinv, const T* tex, const T* axis, const T* sharp, const T* amp, const T* rough, const T* target,
                          const float* cov_clip = nullptr) {
    int nv = s->views > 1 ? s->views : 1;
    size_t P = (size_t)s->W * s->H * nv;
    int C = s->C;
    size_t ntex = level_offset(s->tw, s->th, C, s->L);
    auto sz = [](std::vector<T>& v, size_t n) { v.assign(n, T(0)); };
    sz(w.clip, (size_t)nv * s->V * 4); sz(w.rast, P * 4); sz(w.db, P * 4); sz(w.uvp, P * 2); sz(w.uvDA, P * 4);
    sz(w.nrm, P * 3); sz(w.view, P * 3); sz(w.diff, P * C); sz(w.sAxis, P * 3); sz(w.sSharp, P); sz(w.sAmp, P);
    sz(w.color, P * C); sz(w.denv, P * C); sz(w.senv, P * C); sz(w.img, P * C); sz(w.gimg, P * C); sz(w.gcolor, P * C);
    sz(w.gdiff, P * C); sz(w.gtex, ntex); sz(w.guv, P * 2); sz(w.guvDA, P * 4); sz(w.grast, P * 2); sz(w.gclip, (size_t)nv * s->V * 4);
    sz(w.gpos, (size_t)s->V * 3); sz(w.gaxis, (size_t)s->K * 3); sz(w.gsharp, s->K); sz(w.gamp, (size_t)s->K * C);
    // forward
    project_fwd<T>(s->V, nv, pos, mat, w.clip.data());
    raster_fwd<T>(s->W, s->H, nv, s->V, w.clip.data(), s->T, s->idx, w.rast.data(), w.db.data(), cov_clip);
    interp_fwd<T>(s->W, s->H, nv, 2, w.rast.data(), w.db.data(), uv, s->idx, w.uvp.data(), w.uvDA.data());
    normal_axis<T>(s->W, s->H, nv, w.rast.data(), normals, s->idx, rot, w.nrm.data());
    view_axis<T>(s->W, s->H, nv, pvinv, w.view.data());
    tex_fwd<T>(s->W, s->H, nv, s->tw, s->th, C, s->L, tex, w.rast.data(), w.uvp.data(), w.uvDA.data(), w.diff.data());
    sg_spec_fwd<T>(s->W, s->H, nv, w.rast.data(), w.nrm.data(), w.view.data(), rough, (T)s->ior, w.sAxis.data(), w.sSharp.data(), w.sAmp.data());
    sg_fwd<T>(s->W, s->H, nv, C, s->K, w.rast.data(), axis, sharp, amp, w.nrm.data(), w.diff.data(), w.sAxis.data(), w.sSharp.data(), w.sAmp.data(), w.color.data(), w.denv.data(), w.senv.data());
    aa_fwd<T>(s->W, s->H, nv, C, s->V, w.clip.data(), s->idx, w.rast.data(), w.color.data(), w.img.data());
    double loss = mse<T>(P * C, w.img.data(), target, w.gimg.data());
    // backward
    aa_bwd<T>(s->W, s->H, nv, C, s->V, w.clip.data(), s->idx, w.rast.data(), w.color.data(), w.gimg.data(), w.gcolor.data(), w.gclip.data());
    sg_bwd<T>(s->W, s->H, nv, C, s->K, w.rast.data(), axis, sharp, amp, w.nrm.data(), w.diff.data(), w.sAxis.data(), w.sSharp.data(), w.sAmp.data(),
              w.gcolor.data(), w.gdiff.data(), w.gaxis.data(), w.gsharp.data(), w.gamp.data(), nullptr, nullptr, nullptr, nullptr);
    tex_bwd<T>(s->W, s->H, nv, s->tw, s->th, C, s->L, tex, w.rast.data(), w.uvp.data(), w.uvDA.data(), w.gdiff.data(), w.gtex.data(), w.guv.data(), w.guvDA.data(), 0);
    interp_bwd<T>(s->W, s->H, nv, 2, w.rast.data(), uv, s->idx, w.guv.data(), nullptr, w.grast.data(), s->V);
    raster_bwd<T>(s->W, s->H, nv, s->V, w.clip.data(), s->idx, w.rast.data(), w.grast.data(), w.gclip.data());
    project_bwd<T>(s->V, nv, mat, w.gclip.data(), w.gpos.data());
    return loss;
}

extern "C" void* orc_work_create(void) { return new OrcWork(); }
extern "C" void orc_work_destroy(void* w) { delete (OrcWork*)w; }
extern "C" void* orc_work64_create(void) { return new FitWork<double>(); }
extern "C" void orc_work64_destroy(void* w) { delete (FitWork<double>*)w; }

extern "C" void orc_fit_step_f32(OrcScene* s, void* work) {
    OrcWork& w = *(OrcWork*)work;
    int C = s->C;
    int nv = s->views > 1 ? s->views : 1;
    size_t P = (size_t)s->W * s->H * nv;
    s->loss = fit_fwd_bwd<float>(s, w, s->pos, s->normals, s->uv, s->mat, s->rot, s->pvinv, s->tex, s->axis, s->sharp, s->amp, s->rough, s->target);
    // update
    s->b1t *= s->b1; s->b2t *= s->b2;
    mip_fold<float>(s->tw, s->th, C, s->L, w.gtex.data(), s->fold_quarter);
    adam<float>((size_t)s->tw * s->th * C, s->tex, w.gtex.data(), s->m_tex, s->v_tex, s->b1, s->b2, s->b1t, s->b2t, s->lr, s->eps);
    clamp_<float>((size_t)s->tw * s->th * C, s->tex, 0.f, 1.f);
    mip_build<float>(s->tw, s->th, C, s->L, s->tex);
    adam<float>((size_t)s->V * 3, s->pos, w.gpos.data(), s->m_pos, s->v_pos, s->b1, s->b2, s->b1t, s->b2t, s->lr, s->eps);
    adam<float>((size_t)s->K * 3, s->axis, w.gaxis.data(), s->m_axis, s->v_axis, s->b1, s->b2, s->b1t, s->b2t, s->lr, s->eps);
    adam<float>((size_t)s->K, s->sharp, w.gsharp.data(), s->m_sharp, s->v_sharp, s->b1, s->b2, s->b1t, s->b2t, s->lr, s->eps);
    adam<float>((size_t)s->K * C, s->amp, w.gamp.data(), s->m_amp, s->v_amp, s->b1, s->b2, s->b1t, s->b2t, s->lr, s->eps);
    if (s->out_image) std::memcpy(s->out_image, w.img.data(), sizeof(float) * P * C);
}

// The same forward and backward in double from the scene's float inputs (no update); the
// texture gradient is folded into level 0 as the step folds it.  Returns the loss.
extern "C" double orc_fit_grads_f64(const OrcScene* s, void* work) {
    FitWork<double>& w = *(FitWork<double>*)work;
    int nv = s->views > 1 ? s->views : 1;
    size_t P = (size_t)s->W * s->H * nv;
    int C = s->C;
    auto cp = [](std::vector<double>& d, const float* src, size_t n) { d.assign(src, src + n); };
    cp(w.in_pos, s->pos, (size_t)s->V * 3); cp(w.in_normals, s->normals, (size_t)s->V * 3); cp(w.in_uv, s->uv, (size_t)s->V * 2);
    cp(w.in_mat, s->mat, (size_t)nv * 16); cp(w.in_rot, s->rot, 16); cp(w.in_pvinv, s->pvinv, (size_t)nv * 9);
    cp(w.in_tex, s->tex, level_offset(s->tw, s->th, C, s->L)); cp(w.in_axis, s->axis, (size_t)s->K * 3);
    cp(w.in_sharp, s->sharp, s->K); cp(w.in_amp, s->amp, (size_t)s->K * C); cp(w.in_rough, s->rough, P);
    cp(w.in_target, s->target, P * C);
    std::vector<float> cov((size_t)nv * s->V * 4);
    project_fwd<float>(s->V, nv, s->pos, s->mat, cov.data());
    double loss = fit_fwd_bwd<double>(s, w, w.in_pos.data(), w.in_normals.data(), w.in_uv.data(), w.in_mat.data(), w.in_rot.data(),
                                      w.in_pvinv.data(), w.in_tex.data(), w.in_axis.data(), w.in_sharp.data(), w.in_amp.data(),
                                      w.in_rough.data(), w.in_target.data(), cov.data());
    mip_fold<double>(s->tw, s->th, C, s->L, w.gtex.data(), s->fold_quarter);
    return loss;
}

// Accessors for the work buffers of the last step (tests compare intermediates).
template <class T>
static const T* work_buffer(FitWork<T>& w, const char* name) {
#define B(x) if (!std::strcmp(name, #x)) return w.x.data();
    B(clip) B(rast) B(db) B(uvp) B(uvDA) B(nrm) B(view) B(diff) B(sAxis) B(sSharp) B(sAmp) B(color) B(denv) B(senv) B(img) B(gimg) B(gcolor)
    B(gdiff) B(guv) B(grast) B(gclip) B(gtex) B(gpos) B(gaxis) B(gsharp) B(gamp)
#undef B
    return nullptr;
}
extern "C" const float* orc_work_buffer(void* work, const char* name) { return work_buffer(*(OrcWork*)work, name); }
extern "C" const double* orc_work64_buffer(void* work, const char* name) { return work_buffer(*(FitWork<double>*)work, name); }
