// tuning.hip -- process-wide kernel schedule parameters (vcl_set_tuning / vcl_get_tuning).
#include "internal.hpp"

vcl_tuning g_vcl_tuning = {16, 256, 16};

thread_local VRecorder* g_vcl_recorder = nullptr;

static bool tuning_ok(const vcl_tuning& t) {
    bool tile = t.raster_tile == 8 || t.raster_tile == 16 || t.raster_tile == 32;
    bool chunk = t.bin_chunk == 64 || t.bin_chunk == 128 || t.bin_chunk == 256;
    bool splat = t.splat_rows == 16 || t.splat_rows == 32 || t.splat_rows == 64;
    return tile && chunk && splat;
}

extern "C" int vcl_set_tuning(const vcl_tuning* t) {
    VCL_REQUIRE(t && tuning_ok(*t));
    g_vcl_tuning = *t;
    return VCL_OK;
}

extern "C" int vcl_get_tuning(vcl_tuning* t) {
    VCL_REQUIRE(t);
    *t = g_vcl_tuning;
    return VCL_OK;
}
