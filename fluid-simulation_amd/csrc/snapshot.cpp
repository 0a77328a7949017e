// .sphs snapshots: a 76-byte header {char magic[4] = "SPHS"; u32 version = 1;
// char app_name[64] (NUL-terminated); u32 fluid_n} followed by float4
// positions[fluid_n] and float4 velocities[fluid_n], fluid particles only.
// The layout is the reference's (src/cuda/SPH/snapshot.cuh:9-51) so files
// move between the two; the checks and messages follow Snapshot::save/load
// (snapshot.cu:10-106).  Host code only: no device is needed.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/wcsph.h"

namespace {

thread_local std::string g_snap_error;

struct Header {
    char magic[4];
    uint32_t version;
    char app_name[64];
    uint32_t fluid_n;
};
static_assert(sizeof(Header) == 76, "snapshot header is 76 bytes");

int snap_fail(const std::string& msg) {
    g_snap_error = msg;
    return WCSPH_ERR_INVALID_ARGUMENT;
}

int read_header(std::FILE* f, const char* path, Header& h) {
    if (std::fread(&h, sizeof(h), 1, f) != 1) return snap_fail(std::string("snapshot: header read failed for ") + path);
    if (std::memcmp(h.magic, "SPHS", 4) != 0) return snap_fail(std::string("snapshot: bad magic in ") + path);
    if (h.version != 1)
        return snap_fail("snapshot: version mismatch (file=" + std::to_string(h.version) + ", expected=1)");
    h.app_name[63] = '\0';  // guard against corrupt files
    // a count the body cannot hold is refused before anyone sizes buffers by it
    const long here = std::ftell(f);
    if (here < 0 || std::fseek(f, 0, SEEK_END) != 0) return snap_fail(std::string("snapshot: cannot seek in ") + path);
    const long end = std::ftell(f);
    if (end < 0 || std::fseek(f, here, SEEK_SET) != 0)
        return snap_fail(std::string("snapshot: cannot seek in ") + path);
    if ((uint64_t)(end - here) < 32ull * h.fluid_n)
        return snap_fail(std::string("snapshot: body read failed for ") + path + " (" + std::to_string(h.fluid_n) +
                         " particles need " + std::to_string(32ull * h.fluid_n) + " bytes, the file holds " +
                         std::to_string(end - here) + ")");
    return WCSPH_OK;
}

}  // namespace

extern "C" {

const char* wcsph_snapshot_error(void) { return g_snap_error.c_str(); }

int wcsph_snapshot_info(const char* path, char* app_name, size_t app_name_len, uint32_t* fluid_n) {
    g_snap_error.clear();
    if (!path) return snap_fail("snapshot: NULL path");
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return snap_fail(std::string("snapshot: cannot open ") + path);
    Header h{};
    const int rc = read_header(f, path, h);
    std::fclose(f);
    if (rc) return rc;
    if (app_name && app_name_len) {
        std::strncpy(app_name, h.app_name, app_name_len - 1);
        app_name[app_name_len - 1] = '\0';
    }
    if (fluid_n) *fluid_n = h.fluid_n;
    return WCSPH_OK;
}

int wcsph_snapshot_read(const char* path, const char* app_name, uint32_t fluid_n, float* pos, float* vel) {
    g_snap_error.clear();
    if (!path || !pos || !vel) return snap_fail("snapshot: NULL argument");
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return snap_fail(std::string("snapshot: cannot open ") + path);
    Header h{};
    int rc = read_header(f, path, h);
    if (!rc && app_name && std::strcmp(app_name, h.app_name) != 0)
        rc = snap_fail(std::string("snapshot: app_name mismatch (file='") + h.app_name + "', expected='" + app_name + "')");
    if (!rc && h.fluid_n != fluid_n)
        rc = snap_fail("snapshot: fluid count mismatch (file=" + std::to_string(h.fluid_n) +
                       ", expected=" + std::to_string(fluid_n) + ")");
    if (!rc && (std::fread(pos, 16, fluid_n, f) != fluid_n || std::fread(vel, 16, fluid_n, f) != fluid_n))
        rc = snap_fail(std::string("snapshot: body read failed for ") + path);
    std::fclose(f);
    return rc;
}

int wcsph_snapshot_write(const char* path, const char* app_name, uint32_t fluid_n, const float* pos,
                         const float* vel) {
    g_snap_error.clear();
    if (!path || !app_name || (fluid_n && (!pos || !vel))) return snap_fail("snapshot: NULL argument");
    if (std::strlen(app_name) >= 64) return snap_fail("snapshot: app_name too long (max 63 chars)");
    Header h{};
    std::memcpy(h.magic, "SPHS", 4);
    h.version = 1;
    std::strncpy(h.app_name, app_name, 63);
    h.fluid_n = fluid_n;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return snap_fail(std::string("snapshot: cannot open ") + path);
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    ok = ok && std::fwrite(pos, 16, fluid_n, f) == fluid_n && std::fwrite(vel, 16, fluid_n, f) == fluid_n;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return snap_fail(std::string("snapshot: write failed for ") + path);
    return WCSPH_OK;
}

}  // extern "C"
