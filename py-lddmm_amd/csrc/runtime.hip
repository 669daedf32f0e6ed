// Library runtime: error text, per-device workspace, occupancy cache,
// host-side adjacency builder.
#include "common.hpp"

#include <cmath>
#include <cstdarg>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace lddmm {

namespace {
thread_local std::string g_last_error;

struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mutex;
std::unordered_map<int, Workspace> g_ws;

std::mutex g_info_mutex;
std::unordered_map<int, DeviceInfo> g_info;

std::mutex g_occ_mutex;
std::unordered_map<const void*, int> g_occ;
}  // namespace

void set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
    set_error("HIP error %d (%s) in %s at %s:%d", int(e), hipGetErrorString(e), what, file, line);
    return LDDMM_EHIP;
}

int device_info(DeviceInfo* out) {
    int dev = 0;
    LDDMM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(g_info_mutex);
    auto it = g_info.find(dev);
    if (it == g_info.end()) {
        DeviceInfo d;
        d.device = dev;
        LDDMM_HIP(hipDeviceGetAttribute(&d.num_cu, hipDeviceAttributeMultiprocessorCount, dev));
        it = g_info.emplace(dev, d).first;
    }
    *out = it->second;
    return LDDMM_OK;
}

int workspace(size_t bytes, hipStream_t stream, void** ptr) {
    int dev = 0;
    LDDMM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(g_ws_mutex);
    Workspace& w = g_ws[dev];
    if (bytes > w.bytes) {
        if (w.ptr) {
            // the previous block may still be read by work queued on `stream`
            LDDMM_HIP(hipStreamSynchronize(stream));
            LDDMM_HIP(hipFree(w.ptr));
            w.ptr = nullptr;
            w.bytes = 0;
        }
        size_t grow = bytes + bytes / 4;
        grow = (grow + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
        LDDMM_HIP(hipMalloc(&w.ptr, grow));
        w.bytes = grow;
    }
    *ptr = w.ptr;
    return LDDMM_OK;
}

int release_workspace() {
    int dev = 0;
    LDDMM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(g_ws_mutex);
    auto it = g_ws.find(dev);
    if (it != g_ws.end() && it->second.ptr) {
        LDDMM_HIP(hipDeviceSynchronize());
        LDDMM_HIP(hipFree(it->second.ptr));
        g_ws.erase(it);
    }
    return LDDMM_OK;
}

int blocks_per_cu(const void* func, int block, size_t dyn_lds) {
    std::lock_guard<std::mutex> g(g_occ_mutex);
    auto it = g_occ.find(func);
    if (it != g_occ.end()) return it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, func, block, dyn_lds) != hipSuccess || nb < 1)
        nb = 1;
    g_occ[func] = nb;
    return nb;
}

// ---------------------------------------------------------------------------
// Pair-kernel timing: event pairs are recorded around each launch and read
// back (synchronously) by lddmm_profile_read.
// ---------------------------------------------------------------------------
namespace {
struct ProfState {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> open;
    std::vector<hipEvent_t> pool;
};
std::mutex g_prof_mutex;
ProfState g_prof;

hipEvent_t take_event() {
    if (!g_prof.pool.empty()) {
        hipEvent_t e = g_prof.pool.back();
        g_prof.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

thread_local bool t_prof_off = false;

bool profiling() { return g_prof.on && !t_prof_off; }

ProfileOff::ProfileOff() : prev(t_prof_off) { t_prof_off = true; }
ProfileOff::~ProfileOff() { t_prof_off = prev; }

int profile_begin(hipStream_t st) {
    std::lock_guard<std::mutex> g(g_prof_mutex);
    hipEvent_t a = take_event(), b = take_event();
    if (!a || !b) { set_error("hipEventCreate failed"); return LDDMM_EHIP; }
    LDDMM_HIP(hipEventRecord(a, st));
    g_prof.open.push_back({a, b});
    return LDDMM_OK;
}

int profile_end(hipStream_t st) {
    std::lock_guard<std::mutex> g(g_prof_mutex);
    if (g_prof.open.empty()) return LDDMM_OK;
    LDDMM_HIP(hipEventRecord(g_prof.open.back().second, st));
    return LDDMM_OK;
}

int profile_read(double* ms_each, int64_t cap, int64_t* launches) {
    std::lock_guard<std::mutex> g(g_prof_mutex);
    int64_t i = 0;
    for (auto& ev : g_prof.open) {
        LDDMM_HIP(hipEventSynchronize(ev.second));
        float t = 0.f;
        LDDMM_HIP(hipEventElapsedTime(&t, ev.first, ev.second));
        if (ms_each && i < cap) ms_each[i] = t;
        ++i;
        g_prof.pool.push_back(ev.first);
        g_prof.pool.push_back(ev.second);
    }
    if (launches) *launches = i;
    g_prof.open.clear();
    return LDDMM_OK;
}

int validate_kernel(const lddmm_kernel* kd, unsigned flags) {
    LDDMM_REQUIRE(kd != nullptr, "null kernel description");
    if (kd->affine != LDDMM_AFFINE_NONE) {
        if (!(flags & kValAffine)) {
            set_error("affine kernel terms are applied by the shooting entry points only (got affine = %d)",
                      kd->affine);
            return LDDMM_EUNSUPPORTED;
        }
        LDDMM_REQUIRE(kd->affine == LDDMM_AFFINE_FULL || kd->affine == LDDMM_AFFINE_EUCLIDEAN,
                      "bad affine kind %d", kd->affine);
        LDDMM_REQUIRE(std::isfinite(kd->w1) && std::isfinite(kd->w2), "affine weights must be finite");
        for (int i = 0; i < 3; ++i)
            LDDMM_REQUIRE(std::isfinite(kd->center[i]), "affine center must be finite");
    }
    if (kd->kind == LDDMM_KERNEL_LAPLACIAN) {
        if (flags & kValDataTerm) {
            set_error("the fused varifold data terms take the gauss kernel only (got laplacian)");
            return LDDMM_EUNSUPPORTED;
        }
        LDDMM_REQUIRE(kd->order >= 0 && kd->order <= 4, "laplacian order %d outside [0, 4]", kd->order);
    } else if (kd->kind != LDDMM_KERNEL_GAUSS) {
        set_error("kernel kind %d not supported (gauss = 0, laplacian = 1)", kd->kind);
        return LDDMM_EUNSUPPORTED;
    }
    LDDMM_REQUIRE(kd->nscale >= 1 && kd->nscale <= LDDMM_MAX_SCALES,
                  "nscale %d outside [1, %d]", kd->nscale, LDDMM_MAX_SCALES);
    for (int i = 0; i < kd->nscale; ++i)
        LDDMM_REQUIRE(std::isfinite(kd->sigma[i]) && kd->sigma[i] > 0, "sigma[%d] must be > 0", i);
    LDDMM_REQUIRE(kd->prec == LDDMM_PREC_F64 || kd->prec == LDDMM_PREC_F32C, "bad precision %d", kd->prec);
    return LDDMM_OK;
}

}  // namespace lddmm

extern "C" {

const char* lddmm_last_error(void) { return lddmm::g_last_error.c_str(); }

const char* lddmm_version(void) { return "lddmm_amd 0.1 (gfx950)"; }

int lddmm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int lddmm_release_workspace(void) { return lddmm::release_workspace(); }

int lddmm_profile_enable(int32_t on) {
    lddmm::g_prof.on = on != 0;
    return LDDMM_OK;
}

int lddmm_profile_read(double* ms_each, int64_t cap, int64_t* launches) {
    return lddmm::profile_read(ms_each, cap, launches);
}

}  // extern "C"

namespace lddmm {
namespace {
// CSR vertex -> (element*K + corner) adjacency for elements of K vertices.
// Elements are visited in increasing order, so each vertex's list is sorted
// by (element, corner): the gathers sum in the same order on every call.
int build_adjacency(const int32_t* el, int64_t ne, int K, int64_t nv, int32_t* adj_ptr, int32_t* adj_idx,
                    const char* what) {
    LDDMM_REQUIRE(ne >= 0 && nv >= 0, "%s: negative sizes", what);
    LDDMM_REQUIRE(K * ne < (int64_t(1) << 31), "%s: too many elements for int32 adjacency", what);
    LDDMM_REQUIRE(ne == 0 || (el && adj_idx), "%s: null elements / adj_idx", what);
    LDDMM_REQUIRE(adj_ptr != nullptr, "%s: null adj_ptr", what);
    std::vector<int32_t> cnt(size_t(nv) + 1, 0);
    for (int64_t f = 0; f < ne; ++f)
        for (int c = 0; c < K; ++c) {
            int32_t v = el[K * f + c];
            LDDMM_REQUIRE(v >= 0 && v < nv, "%s: element %lld references vertex %d outside [0, %lld)", what,
                          (long long)f, v, (long long)nv);
            cnt[size_t(v) + 1]++;
        }
    for (int64_t v = 0; v < nv; ++v) cnt[size_t(v) + 1] += cnt[size_t(v)];
    for (int64_t v = 0; v <= nv; ++v) adj_ptr[v] = cnt[size_t(v)];
    for (int64_t f = 0; f < ne; ++f)
        for (int c = 0; c < K; ++c) {
            int32_t v = el[K * f + c];
            adj_idx[cnt[size_t(v)]++] = int32_t(K * f + c);
        }
    return LDDMM_OK;
}
}  // namespace
}  // namespace lddmm

extern "C" {

int lddmm_surface_adjacency(const int32_t* faces, int64_t nf, int64_t nv, int32_t* adj_ptr,
                            int32_t* adj_idx) {
    return lddmm::build_adjacency(faces, nf, 3, nv, adj_ptr, adj_idx, "surface_adjacency");
}

int lddmm_curve_adjacency(const int32_t* segments, int64_t ns, int64_t nv, int32_t* adj_ptr,
                          int32_t* adj_idx) {
    return lddmm::build_adjacency(segments, ns, 2, nv, adj_ptr, adj_idx, "curve_adjacency");
}

}  // extern "C"
