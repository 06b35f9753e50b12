// gqis_runtime.cpp -- C ABI of the sweep integrator: hiprtc compile cache, module launch,
// events and the kernel status word.  Declarations and reference citations: include/gqis_hip.h.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <unistd.h>

#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gqis_hip.h"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

int hip_fail(hipError_t err, const char* what) {
    return fail(GQ_E_HIP, std::string(what) + ": " + hipGetErrorString(err));
}

// Scratch of gq_launch_sweep_noalloc: one array per GPU for the life of the process.  Host
// code never uses this name as an address (on the host it is only a stand-in); every use goes
// through the per-device pointer from hipGetSymbolAddress kept in gq_ctx::scratch.
static __device__ __align__(256) unsigned char scratch[GQ_SCRATCH_BYTES];

// Work cursor (and the caller's status word, when given) back to zero before a chain.
__global__ void gq_scratch_reset(unsigned int* cursor, unsigned int* status) {
    if (threadIdx.x == 0) {
        *cursor = 0u;
        if (status) *status = 0u;
    }
}

// FNV-1a 64-bit over the source, options and architecture: the module cache key.
unsigned long long fnv1a(const std::string& text, unsigned long long h = 1469598103934665603ull) {
    for (unsigned char c : text) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

struct Compiled {
    std::vector<char> code;
    std::string log;
    int rc = GQ_OK;
};

// Code-object files of the on-disk cache: "<dir>/<arch>-<key>-<check>.hsaco", where key and
// check are two independent 64-bit hashes of the same key text (compiler version, architecture,
// options, source).  Written to a temporary name and renamed, so a reader never sees a partial
// file; a file that fails to load is ignored and rebuilt.
bool read_file(const std::string& path, std::vector<char>* out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.seekg(0, std::ios::end);
    std::streamoff n = f.tellg();
    if (n <= 0) return false;
    out->resize((size_t)n);
    f.seekg(0, std::ios::beg);
    return (bool)f.read(out->data(), n);
}

void write_file_atomic(const std::string& path, const std::vector<char>& data) {
    std::string tmp = path + "." + std::to_string((long long)getpid()) + ".tmp";
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) return;
        f.write(data.data(), (std::streamsize)data.size());
        if (!f) {
            f.close();
            std::remove(tmp.c_str());
            return;
        }
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
}

Compiled hiprtc_compile(const char* source, const std::string& arch, const char* const* options,
                        int num_options) {
    Compiled out;
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, source, "gqis_kernel.hip", 0, nullptr, nullptr) !=
        HIPRTC_SUCCESS) {
        out.rc = GQ_E_COMPILE;
        out.log = "hiprtcCreateProgram failed";
        return out;
    }
    std::vector<std::string> opts;
    opts.push_back("--offload-arch=" + arch);
    for (int i = 0; i < num_options; ++i) opts.emplace_back(options[i]);
    std::vector<const char*> argv;
    for (auto& o : opts) argv.push_back(o.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)argv.size(), argv.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    if (log_size > 1) {
        out.log.resize(log_size);
        hiprtcGetProgramLog(prog, &out.log[0]);
    }
    if (r != HIPRTC_SUCCESS) {
        out.rc = GQ_E_COMPILE;
        out.log = std::string("hiprtc: ") + hiprtcGetErrorString(r) + "\n" + out.log;
        hiprtcDestroyProgram(&prog);
        return out;
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(prog, &code_size);
    out.code.resize(code_size);
    hiprtcGetCode(prog, out.code.data());
    hiprtcDestroyProgram(&prog);
    return out;
}

}  // namespace

struct gq_module {
    gq_ctx* ctx = nullptr;
    hipModule_t module = nullptr;
    std::map<std::string, hipFunction_t> functions;
    std::mutex mu;
    int refs = 1;                    // the context's cache holds one; callers hold the others
    unsigned long long last_use = 0;  // context clock of the last compile call that returned it
};

struct gq_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

struct gq_ctx {
    int device = 0;
    std::string arch;
    int compute_units = 0;
    std::string cache_dir;   // on-disk code-object cache; empty = disabled
    unsigned char* scratch = nullptr;   // this device's copy of `scratch`
    std::mutex mu;
    std::map<unsigned long long, gq_module*> modules;
    unsigned long long clock = 0;
};

namespace {

// Loaded modules a context keeps for reuse.  Beyond this many, modules that no caller holds
// any more (only the cache's reference left) are unloaded, least recently used first.
constexpr size_t kModuleCacheMax = 128;

// With ctx->mu held.
void evict_idle_modules(gq_ctx* ctx) {
    while (ctx->modules.size() > kModuleCacheMax) {
        auto victim = ctx->modules.end();
        for (auto it = ctx->modules.begin(); it != ctx->modules.end(); ++it) {
            if (it->second->refs == 1 &&
                (victim == ctx->modules.end() || it->second->last_use < victim->second->last_use))
                victim = it;
        }
        if (victim == ctx->modules.end()) return;   // every module is still held by a caller
        if (victim->second->module) (void)hipModuleUnload(victim->second->module);
        delete victim->second;
        ctx->modules.erase(victim);
    }
}

}  // namespace

extern "C" {

const char* gq_last_error(void) { return g_error.c_str(); }

unsigned long long gq_sweep_args_size(void) { return sizeof(gq_sweep_args); }

int gq_ctx_create(int device_ordinal, gq_ctx** out) {
    if (!out) return fail(GQ_E_INVALID, "gq_ctx_create: null output pointer");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device_ordinal < 0 || device_ordinal >= count)
        return fail(GQ_E_INVALID, "gq_ctx_create: device ordinal out of range");
    e = hipSetDevice(device_ordinal);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device_ordinal);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    // the device address of this GPU's scratch array (the device is current)
    void* scratch_ptr = nullptr;
    e = hipGetSymbolAddress(&scratch_ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    gq_ctx* ctx = new gq_ctx();
    ctx->device = device_ordinal;
    ctx->scratch = (unsigned char*)scratch_ptr;
    std::string arch = prop.gcnArchName;
    ctx->arch = arch.substr(0, arch.find(':'));   // "gfx950:sramecc+:xnack-" -> "gfx950"
    ctx->compute_units = prop.multiProcessorCount;
    *out = ctx;
    return GQ_OK;
}

void gq_ctx_destroy(gq_ctx* ctx) {
    if (!ctx) return;
    for (auto& kv : ctx->modules) {
        if (kv.second->module) (void)hipModuleUnload(kv.second->module);
        delete kv.second;
    }
    delete ctx;
}

int gq_ctx_info(gq_ctx* ctx, char* arch, int arch_len, int* compute_units) {
    if (!ctx) return fail(GQ_E_INVALID, "gq_ctx_info: null context");
    if (arch && arch_len > 0) {
        std::snprintf(arch, (size_t)arch_len, "%s", ctx->arch.c_str());
    }
    if (compute_units) *compute_units = ctx->compute_units;
    return GQ_OK;
}

int gq_ctx_set_cache_dir(gq_ctx* ctx, const char* dir) {
    if (!ctx) return fail(GQ_E_INVALID, "gq_ctx_set_cache_dir: null context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->cache_dir = dir ? dir : "";
    while (ctx->cache_dir.size() > 1 && ctx->cache_dir.back() == '/') ctx->cache_dir.pop_back();
    return GQ_OK;
}

int gq_compile_check(const char* source, const char* arch, const char* const* options,
                     int num_options, long long* code_size) {
    if (!source || !arch) return fail(GQ_E_INVALID, "gq_compile_check: null source or arch");
    Compiled c = hiprtc_compile(source, arch, options, num_options);
    if (c.rc != GQ_OK) return fail(c.rc, c.log);
    if (code_size) *code_size = (long long)c.code.size();
    return GQ_OK;
}

int gq_module_compile(gq_ctx* ctx, const char* source, const char* const* options, int num_options,
                      gq_module** out, int* cached_out, double* build_seconds) {
    if (!ctx || !source || !out) return fail(GQ_E_INVALID, "gq_module_compile: null argument");
    auto t0 = std::chrono::steady_clock::now();
    int rtc_major = 0, rtc_minor = 0;
    (void)hiprtcVersion(&rtc_major, &rtc_minor);
    std::string key_text = "hiprtc " + std::to_string(rtc_major) + "." +
                           std::to_string(rtc_minor) + '\0' + ctx->arch;
    key_text.push_back('\0');
    for (int i = 0; i < num_options; ++i) {
        key_text += options[i];
        key_text.push_back('\0');
    }
    unsigned long long key = fnv1a(source, fnv1a(key_text));
    std::string dir;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        auto it = ctx->modules.find(key);
        if (it != ctx->modules.end()) {
            it->second->refs++;
            it->second->last_use = ++ctx->clock;
            *out = it->second;
            if (cached_out) *cached_out = 1;
            if (build_seconds) *build_seconds = 0.0;
            return GQ_OK;
        }
        dir = ctx->cache_dir;
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipModule_t mod = nullptr;
    int cached = 0;
    std::string path;
    if (!dir.empty()) {
        // second hash: FNV-1a with another offset basis over the same text
        unsigned long long check = fnv1a(source, fnv1a(key_text, 0x84222325cbf29ce4ull));
        char name[96];
        std::snprintf(name, sizeof(name), "/%s-%016llx-%016llx.hsaco", ctx->arch.c_str(), key,
                      check);
        path = dir + name;
        std::vector<char> code;
        if (read_file(path, &code) && hipModuleLoadData(&mod, code.data()) == hipSuccess) {
            cached = 2;
        } else {
            mod = nullptr;
        }
    }
    if (!mod) {
        Compiled c = hiprtc_compile(source, ctx->arch, options, num_options);
        if (c.rc != GQ_OK) return fail(c.rc, c.log);
        e = hipModuleLoadData(&mod, c.code.data());
        if (e != hipSuccess) return hip_fail(e, "hipModuleLoadData");
        if (!path.empty()) write_file_atomic(path, c.code);
    }
    gq_module* m = nullptr;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        auto it = ctx->modules.find(key);
        if (it != ctx->modules.end()) {
            // another thread compiled the same key meanwhile: keep its module, drop this one
            m = it->second;
            m->refs++;
            (void)hipModuleUnload(mod);
        } else {
            m = new gq_module();
            m->ctx = ctx;
            m->module = mod;
            m->refs = 2;   // one for the cache, one for the caller
            ctx->modules[key] = m;
        }
        m->last_use = ++ctx->clock;
        evict_idle_modules(ctx);
    }
    *out = m;
    if (cached_out) *cached_out = cached;
    if (build_seconds) {
        *build_seconds =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return GQ_OK;
}

static int get_function(gq_module* m, const char* name, hipFunction_t* fn) {
    std::lock_guard<std::mutex> lock(m->mu);
    auto it = m->functions.find(name);
    if (it != m->functions.end()) {
        *fn = it->second;
        return GQ_OK;
    }
    hipError_t e = hipModuleGetFunction(fn, m->module, name);
    if (e != hipSuccess)
        return fail(GQ_E_NOT_FOUND, std::string("kernel not found in module: ") + name);
    m->functions[name] = *fn;
    return GQ_OK;
}

int gq_launch_sweep(gq_module* module, const char* name, unsigned int grid, unsigned int block,
                    const gq_sweep_args* args, void* stream) {
    if (!module || !name || !args) return fail(GQ_E_INVALID, "gq_launch_sweep: null argument");
    if (grid == 0 || block == 0 || block > 1024)
        return fail(GQ_E_INVALID, "gq_launch_sweep: bad launch shape");
    hipFunction_t fn;
    int rc = get_function(module, name, &fn);
    if (rc != GQ_OK) return rc;
    gq_sweep_args local = *args;
    void* params[] = {&local};
    hipError_t e = hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, 0, (hipStream_t)stream,
                                         params, nullptr);
    if (e != hipSuccess) return hip_fail(e, "hipModuleLaunchKernel");
    return GQ_OK;
}

int gq_graph_create(gq_module* module, const gq_launch_desc* launches, int count,
                    const gq_sweep_args* args, gq_graph** out) {
    if (!module || !launches || count <= 0 || !args || !out)
        return fail(GQ_E_INVALID, "gq_graph_create: null argument or empty launch list");
    std::vector<hipFunction_t> fns((size_t)count);
    for (int i = 0; i < count; ++i) {
        if (!launches[i].name || launches[i].grid == 0 || launches[i].block == 0 ||
            launches[i].block > 1024)
            return fail(GQ_E_INVALID, "gq_graph_create: bad launch description");
        int rc = get_function(module, launches[i].name, &fns[(size_t)i]);
        if (rc != GQ_OK) return rc;
    }
    hipError_t e = hipSetDevice(module->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    // capture on a private stream, so that no other work of the process is recorded
    hipStream_t s = nullptr;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreateWithFlags");
    gq_graph* g = new gq_graph();
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        gq_sweep_args local = *args;
        void* params[] = {&local};
        for (int i = 0; i < count && e == hipSuccess; ++i) {
            e = hipModuleLaunchKernel(fns[(size_t)i], launches[i].grid, 1, 1, launches[i].block,
                                      1, 1, 0, s, params, nullptr);
        }
        hipGraph_t graph = nullptr;
        hipError_t e_end = hipStreamEndCapture(s, &graph);
        if (e == hipSuccess) e = e_end;
        g->graph = graph;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) {
        gq_graph_destroy(g);
        return hip_fail(e, "gq_graph_create");
    }
    *out = g;
    return GQ_OK;
}

int gq_graph_launch(gq_graph* graph, void* stream) {
    if (!graph || !graph->exec) return fail(GQ_E_INVALID, "gq_graph_launch: null graph");
    hipError_t e = hipGraphLaunch(graph->exec, (hipStream_t)stream);
    return e == hipSuccess ? GQ_OK : hip_fail(e, "hipGraphLaunch");
}

int gq_launch_sweep_noalloc(gq_module* module, const gq_launch_desc* launches, int count,
                            const gq_sweep_args* args, unsigned long long table_bytes,
                            void* stream) {
    if (!module || !launches || count <= 0 || !args)
        return fail(GQ_E_INVALID, "gq_launch_sweep_noalloc: null argument or empty launch list");
    const unsigned long long table_max = GQ_SCRATCH_BYTES - GQ_SCRATCH_TABLE_OFFSET;
    if (table_bytes > table_max)
        return fail(GQ_E_INVALID, "gq_launch_sweep_noalloc: a time table of " +
                                      std::to_string(table_bytes) + " bytes exceeds the " +
                                      std::to_string(table_max) + "-byte scratch table");
    unsigned char* base = module->ctx->scratch;
    if (!base) return fail(GQ_E_HIP, "gq_launch_sweep_noalloc: no scratch address for the device");
    std::vector<hipFunction_t> fns((size_t)count);
    for (int i = 0; i < count; ++i) {
        if (!launches[i].name || launches[i].grid == 0 || launches[i].block == 0 ||
            launches[i].block > 1024)
            return fail(GQ_E_INVALID, "gq_launch_sweep_noalloc: bad launch description");
        int rc = get_function(module, launches[i].name, &fns[(size_t)i]);
        if (rc != GQ_OK) return rc;
    }
    gq_sweep_args local = *args;
    local.work_cursor = (unsigned int*)base;
    if (table_bytes) local.time_table = base + GQ_SCRATCH_TABLE_OFFSET;
    // the reset kernel runs on the current device: make it the module's for this call
    int previous = 0;
    hipError_t e = hipGetDevice(&previous);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (previous != module->ctx->device) {
        e = hipSetDevice(module->ctx->device);
        if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    }
    gq_scratch_reset<<<1, 64, 0, (hipStream_t)stream>>>(local.work_cursor, local.status);
    e = hipGetLastError();
    void* params[] = {&local};
    for (int i = 0; i < count && e == hipSuccess; ++i) {
        e = hipModuleLaunchKernel(fns[(size_t)i], launches[i].grid, 1, 1, launches[i].block, 1,
                                  1, 0, (hipStream_t)stream, params, nullptr);
    }
    if (previous != module->ctx->device) (void)hipSetDevice(previous);
    return e == hipSuccess ? GQ_OK : hip_fail(e, "gq_launch_sweep_noalloc");
}

void gq_graph_destroy(gq_graph* graph) {
    if (!graph) return;
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
}

int gq_resident_groups(gq_module* module, const char* name, unsigned int block, int* groups) {
    if (!module || !name || !groups) return fail(GQ_E_INVALID, "gq_resident_groups: null argument");
    hipFunction_t fn;
    int rc = get_function(module, name, &fn);
    if (rc != GQ_OK) return rc;
    int per_cu = 0;
    hipError_t e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)block, 0);
    if (e != hipSuccess) return hip_fail(e, "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
    *groups = per_cu * module->ctx->compute_units;
    return GQ_OK;
}

void gq_module_release(gq_module* module) {
    if (!module) return;
    std::lock_guard<std::mutex> lock(module->ctx->mu);
    // the cache keeps its own reference: an unheld module stays loaded for reuse until the
    // cache outgrows kModuleCacheMax (evict_idle_modules) or the context is destroyed
    if (module->refs > 1) module->refs--;
}

int gq_event_create(void** event) {
    if (!event) return fail(GQ_E_INVALID, "gq_event_create: null output");
    hipEvent_t ev;
    hipError_t e = hipEventCreate(&ev);
    if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
    *event = (void*)ev;
    return GQ_OK;
}

int gq_event_record(void* event, void* stream) {
    hipError_t e = hipEventRecord((hipEvent_t)event, (hipStream_t)stream);
    return e == hipSuccess ? GQ_OK : hip_fail(e, "hipEventRecord");
}

int gq_event_elapsed_ms(void* start, void* stop, float* ms) {
    hipError_t e = hipEventSynchronize((hipEvent_t)stop);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    e = hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
    return e == hipSuccess ? GQ_OK : hip_fail(e, "hipEventElapsedTime");
}

int gq_event_destroy(void* event) {
    hipError_t e = hipEventDestroy((hipEvent_t)event);
    return e == hipSuccess ? GQ_OK : hip_fail(e, "hipEventDestroy");
}

int gq_stream_synchronize(void* stream) {
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    return e == hipSuccess ? GQ_OK : hip_fail(e, "hipStreamSynchronize");
}

int gq_status_read(unsigned int* device_status, void* stream, unsigned int* flags, int clear) {
    if (!device_status || !flags) return fail(GQ_E_INVALID, "gq_status_read: null argument");
    unsigned int host = 0;
    hipError_t e = hipMemcpyAsync(&host, device_status, sizeof(host), hipMemcpyDeviceToHost,
                                  (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(status)");
    e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    *flags = host;
    if (clear) {
        e = hipMemsetAsync(device_status, 0, sizeof(unsigned int), (hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync(status)");
    }
    return GQ_OK;
}

}  // extern "C"
