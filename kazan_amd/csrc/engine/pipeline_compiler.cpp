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

namespace {

// Occupancy hints of the two raster kernels (resident waves per SIMD the
// compiler must allow for; 0: its own choice). KAZAN_RASTER_WAVES=n sets
// both to n and turns the occupancy step of compile_pipeline_code off.
// Default 3: binds only the heaviest shaders (an 8-light Phong raster would
// otherwise take a few registers more than 3 waves allow).
struct WaveHints {
    int raster = 3, persist = 3;
};

bool waves_from_env(WaveHints &h) {
    const char *waves = std::getenv("KAZAN_RASTER_WAVES");
    if (!waves) return false;
    h.raster = h.persist = std::max(0, std::min(8, std::atoi(waves)));
    return true;
}

std::string full_source(const FixedState &fs, const std::string &shader_source, const WaveHints &h) {
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
    os << "#define KZ_RASTER_WAVES " << h.raster << "\n";
    os << "#define KZ_PERSIST_WAVES " << h.persist << "\n";
    os << kz_src_device_common << "\n";
    os << "__device__ void kz_sample(const KzTexture &t, float u, float v, float out[4]);\n";
    os << shader_source << "\n";
    os << kz_src_raster_template << "\n";
    return os.str();
}

WaveHints default_hints() {
    WaveHints h;
    waves_from_env(h);
    return h;
}

template <typename T>
T get_le(const std::vector<char> &b, size_t off) {
    T v{};
    std::memcpy(&v, b.data() + off, sizeof v);
    return v;
}

}  // namespace

std::string pipeline_full_source(const FixedState &fs, const std::string &shader_source) {
    return full_source(fs, shader_source, default_hints());
}

bool kernel_resources(const std::vector<char> &elf, const char *kernel, KernelResources &r) {
    if (!is_amdgpu_code_object(elf.data(), elf.size())) return false;
    const size_t n = elf.size();
    const uint64_t shoff = get_le<uint64_t>(elf, 0x28);
    const uint16_t shentsize = get_le<uint16_t>(elf, 0x3A), shnum = get_le<uint16_t>(elf, 0x3C);
    if (shentsize < 64 || shoff > n || (uint64_t)shnum * shentsize > n - shoff) return false;
    auto sh = [&](uint32_t i, size_t field) { return shoff + (size_t)i * shentsize + field; };
    const std::string want = std::string(kernel) + ".kd";
    for (uint32_t i = 0; i < shnum; ++i) {
        if (get_le<uint32_t>(elf, sh(i, 4)) != 2) continue;  // SHT_SYMTAB
        const uint64_t so = get_le<uint64_t>(elf, sh(i, 24)), ss = get_le<uint64_t>(elf, sh(i, 32));
        const uint32_t link = get_le<uint32_t>(elf, sh(i, 40));
        if (link >= shnum || so > n || ss > n - so) return false;
        const uint64_t stro = get_le<uint64_t>(elf, sh(link, 24)), strs = get_le<uint64_t>(elf, sh(link, 32));
        if (stro > n || strs > n - stro) return false;
        for (uint64_t k = 0; k + 24 <= ss; k += 24) {
            const uint32_t name = get_le<uint32_t>(elf, so + k);
            if (name >= strs || strs - name < want.size() + 1 ||
                std::memcmp(elf.data() + stro + name, want.c_str(), want.size() + 1) != 0)
                continue;
            const uint16_t shndx = get_le<uint16_t>(elf, so + k + 6);
            const uint64_t value = get_le<uint64_t>(elf, so + k + 8);
            if (shndx == 0 || shndx >= shnum) return false;
            const uint64_t addr = get_le<uint64_t>(elf, sh(shndx, 16)), off = get_le<uint64_t>(elf, sh(shndx, 24));
            if (value < addr || off + (value - addr) > n || n - (off + (value - addr)) < 64) return false;
            const size_t kd = (size_t)(off + (value - addr));
            r.lds = get_le<uint32_t>(elf, kd + 0);
            r.scratch = get_le<uint32_t>(elf, kd + 4);
            r.vgprs = ((get_le<uint32_t>(elf, kd + 48) & 63u) + 1u) * 8u;
            const uint32_t by_vgpr = 512u / r.vgprs;
            const uint32_t by_lds = r.lds ? (160u * 1024u / r.lds) / 4u : 8u;
            r.waves = std::max(1u, std::min(std::min(by_vgpr, by_lds), 8u));
            return true;
        }
    }
    return false;
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

// the cached code object at `path`, if it is one
bool read_cached(const std::string &path, std::vector<char> &code) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    code.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    return is_amdgpu_code_object(code.data(), code.size());
}

// publish atomically: write a temporary file, then rename
void publish(const std::string &dir, const std::string &path, const std::vector<char> &code) {
    mkdir(dir.c_str(), 0755);
    const std::string tmp = path + ".tmp" + std::to_string(getpid());
    {
        std::ofstream out(tmp, std::ios::binary);
        if (out) out.write(code.data(), (std::streamsize)code.size());
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
}

// hiprtc, no cache
bool rtc_compile(const std::string &src, const char *opt, std::vector<char> &code, std::string &log) {
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
    return true;
}

// hiprtc through the on-disk code-object cache
bool compile_cached(const std::string &src, const char *opt, std::vector<char> &code, std::string &log,
                    bool *cache_hit, bool read_cache) {
    const std::string dir = cache_dir();
    const std::string path = dir + "/" + code_name(src, opt);
    if (cache_hit) *cache_hit = false;
    if (read_cache && read_cached(path, code)) {
        if (cache_hit) *cache_hit = true;
        return true;
    }
    if (!rtc_compile(src, opt, code, log)) return false;
    publish(dir, path, code);
    return true;
}

// The occupancy step for one raster kernel: the hint that asks for one more
// resident wave per SIMD when the kernel's registers are within one
// allocation granule (8 VGPRs) of that level and LDS allows it, else 0.
int next_wave_hint(const KernelResources &r) {
    const uint32_t next = 512u / r.vgprs + 1u;
    const uint32_t by_lds = r.lds ? (160u * 1024u / r.lds) / 4u : 8u;
    if (next > std::min(by_lds, 8u) || r.scratch) return 0;
    const uint32_t budget = (512u / next) & ~7u;
    return r.vgprs <= budget + 8u ? (int)next : 0;
}

}  // namespace

std::string pipeline_code_name(const FixedState &fs, const std::string &shader_source) {
    return code_name(pipeline_full_source(fs, shader_source), opt_level(fs));
}

// The raster kernels are compiled at the default hints first. A kernel whose
// registers are one granule above the next occupancy level is compiled again
// with that level as its hint, and the result kept if the kernel then uses
// no scratch and reaches the level (the raster is latency bound: one more
// resident wave hides more of it). The code object kept is cached under the
// name of the default source, so the choice is made once per pipeline.
bool compile_pipeline_code(const FixedState &fs, const std::string &shader_source, std::vector<char> &code,
                           std::string &log, bool *cache_hit, bool read_cache) {
    WaveHints h;
    const bool forced = waves_from_env(h);
    const std::string src = full_source(fs, shader_source, h);
    const char *opt = opt_level(fs);
    if (forced || !fs.optimize) return compile_cached(src, opt, code, log, cache_hit, read_cache);
    const std::string dir = cache_dir();
    const std::string path = dir + "/" + code_name(src, opt);
    if (cache_hit) *cache_hit = false;
    if (read_cache && read_cached(path, code)) {
        if (cache_hit) *cache_hit = true;
        return true;
    }
    if (!rtc_compile(src, opt, code, log)) return false;
    KernelResources r0, p0;
    WaveHints up = h;
    if (kernel_resources(code, "kz_raster", r0)) up.raster = std::max(h.raster, next_wave_hint(r0));
    if (kernel_resources(code, "kz_raster_persistent", p0)) up.persist = std::max(h.persist, next_wave_hint(p0));
    for (int attempt = 0; attempt < 2 && (up.raster != h.raster || up.persist != h.persist); ++attempt) {
        std::vector<char> alt;
        std::string alt_log;
        if (!rtc_compile(full_source(fs, shader_source, up), opt, alt, alt_log)) break;
        KernelResources r1, p1;
        const bool r_ok = up.raster == h.raster ||
                          (kernel_resources(alt, "kz_raster", r1) && !r1.scratch && (int)r1.waves >= up.raster);
        const bool p_ok = up.persist == h.persist || (kernel_resources(alt, "kz_raster_persistent", p1) &&
                                                      !p1.scratch && (int)p1.waves >= up.persist);
        if (r_ok && p_ok) {
            code.swap(alt);
            break;
        }
        // keep the raise that held, if any, and compile once more with it alone
        if (!r_ok) up.raster = h.raster;
        if (!p_ok) up.persist = h.persist;
    }
    publish(dir, path, code);
    return true;
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
