// pipeline_compiler.cpp — per-pipeline kernel source assembly, hiprtc
// compilation for gfx950 and the on-disk code-object cache.
//
// The cache is the counterpart of the pipeline cache the reference leaves as
// a stub (vkCreatePipelineCache, vulkan-driver/src/api_impl.rs:3979; nil
// pipelineCacheUUID :1489-1492): code objects are keyed by a 64-bit FNV-1a
// hash of the complete kernel source and the compile options, stored as
// <dir>/<hash>.hsaco. The directory is $KAZAN_PIPELINE_CACHE, else
// _pipeline_cache/ beside the driver library.
#include <dlfcn.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>

#include "_embedded_sources.h"
#include "engine.h"

namespace kz {

namespace {

// (the optimisation level, -O3 or -O0 per FixedState::optimize, comes first)
const char *const kCompileOptions[] = {
    "--offload-arch=gfx950",
    "-std=c++17",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
};
constexpr int kNumOptions = sizeof(kCompileOptions) / sizeof(kCompileOptions[0]);

uint64_t fnv1a(const std::string &s) { return fnv1a64(s.data(), s.size()); }

const char *opt_level(const FixedState &fs) { return fs.optimize ? "-O3" : "-O0"; }

std::string cache_dir() {
    if (const char *e = std::getenv("KAZAN_PIPELINE_CACHE")) return e;
    Dl_info info;
    if (dladdr((void *)&cache_dir, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        auto slash = p.rfind('/');
        if (slash != std::string::npos) return p.substr(0, slash) + "/_pipeline_cache";
    }
    return "_pipeline_cache";
}

std::mutex g_compile_mutex;

}  // namespace

std::string pipeline_full_source(const FixedState &fs, const std::string &shader_source) {
    std::ostringstream os;
    os << "// generated pipeline kernel source\n";
    os << "#define KZ_TILE_SHIFT " << fs.tile_shift << "\n";
    os << "#define KZ_SAMPLES " << fs.samples << "\n";
    os << "#define KZ_NV " << fs.num_varyings << "\n";
    os << "#define KZ_CULL_MODE " << fs.cull_mode << "\n";
    os << "#define KZ_FRONT_FACE " << fs.front_face << "\n";
    os << "#define KZ_DEPTH_TEST " << fs.depth_test << "\n";
    os << "#define KZ_DEPTH_WRITE " << fs.depth_write << "\n";
    os << "#define KZ_DEPTH_OP " << fs.depth_op << "\n";
    os << "#define KZ_BLEND " << fs.blend << "\n";
    os << "#define KZ_SRC_COLOR " << fs.src_color << "\n";
    os << "#define KZ_DST_COLOR " << fs.dst_color << "\n";
    os << "#define KZ_COLOR_OP " << fs.color_op << "\n";
    os << "#define KZ_SRC_ALPHA " << fs.src_alpha << "\n";
    os << "#define KZ_DST_ALPHA " << fs.dst_alpha << "\n";
    os << "#define KZ_ALPHA_OP " << fs.alpha_op << "\n";
    os << "#define KZ_WRITE_MASK " << fs.write_mask << "\n";
    // setup records (KAZAN_SETUP_RECORDS=1, read as the engine reads it):
    // compiled out otherwise, so that binning drops what only records need
    const char *rec = std::getenv("KAZAN_SETUP_RECORDS");
    os << "#define KZ_SETUP_RECORDS " << (rec && std::strcmp(rec, "0") != 0 ? 1 : 0) << "\n";
    // KAZAN_RASTER_WAVES=n: ask the compiler for at least n resident raster
    // waves per SIMD (registers beyond the budget spill); 0 = its own
    // choice. Default 3: binds only the heaviest shaders (an 8-light Phong
    // raster would otherwise take a few registers more than 3 waves allow)
    const char *waves = std::getenv("KAZAN_RASTER_WAVES");
    os << "#define KZ_RASTER_WAVES " << (waves ? std::max(0, std::min(8, std::atoi(waves))) : 3) << "\n";
    os << kz_src_device_common << "\n";
    os << "__device__ void kz_sample(const KzTexture &t, float u, float v, float out[4]);\n";
    os << shader_source << "\n";
    os << kz_src_raster_template << "\n";
    return os.str();
}

uint64_t fnv1a64(const void *data, size_t n, uint64_t h) {
    const unsigned char *p = (const unsigned char *)data;
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

bool is_amdgpu_code_object(const char *data, size_t n) {
    // e_ident: 0x7f 'E' 'L' 'F', EI_CLASS = 2 (64-bit), EI_DATA = 1 (LSB);
    // e_machine (offset 18, u16) = EM_AMDGPU (224)
    if (!data || n < 64) return false;
    const unsigned char *p = (const unsigned char *)data;
    return p[0] == 0x7f && p[1] == 'E' && p[2] == 'L' && p[3] == 'F' && p[4] == 2 && p[5] == 1 &&
           (p[18] | (p[19] << 8)) == 224;
}

std::string pipeline_code_dir() { return cache_dir(); }

namespace {

// the cache file name of a complete source compiled at optimisation `opt`
std::string code_name(const std::string &full_source, const char *opt) {
    std::string key_text = full_source;
    key_text += std::string("\n") + opt;
    for (int i = 0; i < kNumOptions; ++i) key_text += std::string("\n") + kCompileOptions[i];
    int rtc_major = 0, rtc_minor = 0;
    hiprtcVersion(&rtc_major, &rtc_minor);
    key_text += "\nhiprtc " + std::to_string(rtc_major) + "." + std::to_string(rtc_minor);
    char name[32];
    std::snprintf(name, sizeof name, "%016llx.hsaco", (unsigned long long)fnv1a(key_text));
    return name;
}

// hiprtc through the on-disk code-object cache
bool compile_cached(const std::string &src, const char *opt, std::vector<char> &code, std::string &log,
                    bool *cache_hit, bool read_cache) {
    const std::string dir = cache_dir();
    const std::string path = dir + "/" + code_name(src, opt);
    if (cache_hit) *cache_hit = false;
    if (read_cache) {
        std::ifstream in(path, std::ios::binary);
        if (in) {
            code.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
            if (is_amdgpu_code_object(code.data(), code.size())) {
                if (cache_hit) *cache_hit = true;
                return true;
            }
        }
    }
    std::lock_guard<std::mutex> lock(g_compile_mutex);
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "kazan_pipeline.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return false;
    }
    const char *options[kNumOptions + 1];
    options[0] = opt;
    for (int i = 0; i < kNumOptions; ++i) options[i + 1] = kCompileOptions[i];
    const hiprtcResult rc = hiprtcCompileProgram(prog, kNumOptions + 1, options);
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    if (n > 1) {
        log.resize(n);
        hiprtcGetProgramLog(prog, &log[0]);
    }
    if (rc != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    // publish atomically: write a temporary file, then rename
    mkdir(dir.c_str(), 0755);
    const std::string tmp = path + ".tmp" + std::to_string(getpid());
    {
        std::ofstream out(tmp, std::ios::binary);
        if (out) out.write(code.data(), (std::streamsize)code.size());
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
    return true;
}

}  // namespace

std::string pipeline_code_name(const FixedState &fs, const std::string &shader_source) {
    return code_name(pipeline_full_source(fs, shader_source), opt_level(fs));
}

bool compile_pipeline_code(const FixedState &fs, const std::string &shader_source, std::vector<char> &code,
                           std::string &log, bool *cache_hit, bool read_cache) {
    return compile_cached(pipeline_full_source(fs, shader_source), opt_level(fs), code, log, cache_hit, read_cache);
}

std::string compute_full_source(const std::string &compute_source) {
    return "// generated compute kernel source\n" + std::string(kz_src_device_common) + "\n" + compute_source;
}

bool compile_compute_code(const std::string &compute_source, bool optimize, std::vector<char> &code, std::string &log,
                          bool *cache_hit) {
    return compile_cached(compute_full_source(compute_source), optimize ? "-O3" : "-O0", code, log, cache_hit, true);
}

bool load_pipeline_kernels(const std::vector<char> &code, PipelineKernels &out, std::string &err) {
    hipError_t e = hipModuleLoadData(&out.module, code.data());
    if (e != hipSuccess) {
        err = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
        return false;
    }
    if ((e = hipModuleGetFunction(&out.bin_count, out.module, "kz_bin_count")) != hipSuccess ||
        (e = hipModuleGetFunction(&out.raster, out.module, "kz_raster")) != hipSuccess ||
        (e = hipModuleGetFunction(&out.raster_persistent, out.module, "kz_raster_persistent")) != hipSuccess) {
        err = std::string("hipModuleGetFunction: ") + hipGetErrorString(e);
        return false;
    }
    // persistent grid: as many workgroups as can be resident at once
    int per_cu = 0, dev = 0, cus = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, out.raster_persistent, 64, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 8;
    out.persistent_blocks = (uint32_t)(std::max(cus, 1) * per_cu);
    return true;
}

PipelineVariants::~PipelineVariants() {
    for (PipelineKernels &k : k_)
        if (k.module) hipModuleUnload(k.module);
}

const PipelineKernels *PipelineVariants::get(uint32_t tile_shift, std::string *err) {
    if (!fits(tile_shift)) {
        if (err) *err = "tile size does not fit the raster workgroup's LDS";
        return nullptr;
    }
    if (!tried_[tile_shift]) {
        tried_[tile_shift] = true;
        FixedState fs = fixed;
        fs.tile_shift = tile_shift;
        bool ok = false;
        std::vector<char> &pre = preloaded_[tile_shift];
        if (!pre.empty()) {  // from a VkPipelineCache
            ok = load_pipeline_kernels(pre, k_[tile_shift], err_[tile_shift]);
            if (!ok) k_[tile_shift] = PipelineKernels{};
            std::vector<char>().swap(pre);
        }
        for (int attempt = 0; !ok && attempt < 2; ++attempt) {
            // the second attempt replaces an on-disk copy that did not load
            std::vector<char> code;
            std::string log;
            bool hit = false;
            if (!compile_pipeline_code(fs, shader_source, code, log, &hit, attempt == 0)) {
                err_[tile_shift] = "kernel compilation failed:\n" + log;
                break;
            }
            ok = load_pipeline_kernels(code, k_[tile_shift], err_[tile_shift]);
            if (!ok) k_[tile_shift] = PipelineKernels{};
            if (!hit) break;
        }
    }
    if (!k_[tile_shift].module) {
        if (err) *err = err_[tile_shift];
        return nullptr;
    }
    return &k_[tile_shift];
}

}  // namespace kz
