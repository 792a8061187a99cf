// Serialization savepoints: the analogue of the reference's serialbox calls
// (advection/src_drivers/advection_CEtimestep_twostep_GPUBC.f90:122-142 and
// 355-375, under SERIALIZE: "!$ser savepoint mysnapshot.run-in", "!$ser data
// uadv,vadv,...").  A serializer is a directory; each field written at a
// savepoint becomes one NumPy .npy file (the array in Fortran order with its
// shape, no pickled objects) and one line of the directory's index
// savepoints.txt:  <savepoint> <field> <file> <dtype> <n1> <n2> <n3>.
// Fields may live in device or host memory; device fields are copied on the
// caller's stream, which is synchronized.  Host code only.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include <sys/stat.h>

#include "../../include/mpdata_hip.h"
#include "mpdata_status.hpp"

using namespace mpdata;

namespace {

std::mutex g_mu;
std::string g_dir;  // the open serializer's directory ("" = none)

bool valid_name(const char* s) {
    if (!s || !*s || strlen(s) > 200) return false;
    for (const char* p = s; *p; p++) {
        const char c = *p;
        if (!((c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') ||
              c == '.' || c == '_' || c == '-'))
            return false;
    }
    return true;
}

// The .npy header of version 1.0: magic, version, header length, then a
// Python dict literal padded with spaces and a newline to a multiple of 64.
std::string npy_header(const char* descr, int n1, int n2, int n3) {
    char dict[256];
    snprintf(dict, sizeof dict, "{'descr': '%s', 'fortran_order': True, 'shape': (%d, %d, %d), }",
             descr, n1, n2, n3);
    std::string d(dict);
    const size_t pre = 10;  // magic (6) + version (2) + length (2)
    size_t total = pre + d.size() + 1;
    const size_t pad = (64 - total % 64) % 64;
    d.append(pad, ' ');
    d.push_back('\n');
    std::string h("\x93NUMPY\x01\x00", 8);
    const unsigned short len = (unsigned short)d.size();
    h.push_back((char)(len & 0xff));
    h.push_back((char)(len >> 8));
    return h + d;
}

template <typename T>
int write_field(const char* savepoint, const char* name, const T* field, int n1, int n2, int n3,
                hipStream_t s, const char* descr, const char* dtype) {
    if (!valid_name(savepoint) || !valid_name(name))
        return fail(MPDATA_EINVAL,
                    "serialize: savepoint and field names are 1-200 characters of [A-Za-z0-9._-]");
    if (!field) return fail(MPDATA_EINVAL, "serialize: null field");
    if (n1 < 1 || n2 < 1 || n3 < 1)
        return fail(MPDATA_EINVAL, "serialize: shape (%d,%d,%d)", n1, n2, n3);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_dir.empty()) return fail(MPDATA_EINVAL, "serialize: no serializer open");
    const size_t n = (size_t)n1 * n2 * n3;
    std::vector<T> host(n);
    hipPointerAttribute_t attr;
    const bool on_device = hipPointerGetAttributes(&attr, field) == hipSuccess &&
                           (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged);
    (void)hipGetLastError();
    if (on_device) {
        HIPCHK(hipMemcpyAsync(host.data(), field, n * sizeof(T), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    } else {
        memcpy(host.data(), field, n * sizeof(T));
    }
    const std::string file = std::string(savepoint) + "__" + name + ".npy";
    const std::string path = g_dir + "/" + file;
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return fail(MPDATA_EINVAL, "serialize: cannot write %s: %s", path.c_str(), strerror(errno));
    const std::string h = npy_header(descr, n1, n2, n3);
    const bool ok = fwrite(h.data(), 1, h.size(), f) == h.size() &&
                    fwrite(host.data(), sizeof(T), n, f) == n;
    if (fclose(f) != 0 || !ok) return fail(MPDATA_EINVAL, "serialize: writing %s failed", path.c_str());
    const std::string index = g_dir + "/savepoints.txt";
    FILE* ix = fopen(index.c_str(), "a");
    if (!ix) return fail(MPDATA_EINVAL, "serialize: cannot append to %s", index.c_str());
    fprintf(ix, "%s %s %s %s %d %d %d\n", savepoint, name, file.c_str(), dtype, n1, n2, n3);
    if (fclose(ix) != 0) return fail(MPDATA_EINVAL, "serialize: writing %s failed", index.c_str());
    return MPDATA_OK;
}

}  // namespace

extern "C" {

int mpdata_serializer_open(const char* dir) {
    if (!dir || !*dir) return fail(MPDATA_EINVAL, "serializer: empty directory name");
    if (mkdir(dir, 0777) != 0 && errno != EEXIST)
        return fail(MPDATA_EINVAL, "serializer: cannot create %s: %s", dir, strerror(errno));
    struct stat st;
    if (stat(dir, &st) != 0 || !S_ISDIR(st.st_mode))
        return fail(MPDATA_EINVAL, "serializer: %s is not a directory", dir);
    std::lock_guard<std::mutex> lk(g_mu);
    g_dir = dir;
    return MPDATA_OK;
}

int mpdata_serializer_close(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_dir.clear();
    return MPDATA_OK;
}

int mpdata_serialize_field_f64(const char* savepoint, const char* name, const double* field,
                               int n1, int n2, int n3, void* stream) {
    return write_field<double>(savepoint, name, field, n1, n2, n3, (hipStream_t)stream, "<f8",
                               "f8");
}

int mpdata_serialize_field_f32(const char* savepoint, const char* name, const float* field,
                               int n1, int n2, int n3, void* stream) {
    return write_field<float>(savepoint, name, field, n1, n2, n3, (hipStream_t)stream, "<f4",
                              "f4");
}

}  // extern "C"
