// backend_abi.cpp — the shader-compiler backend's C ABI (kazan_backend.h):
// compiled code as a loaded hipModule_t whose kernels are looked up by name.
//
// Counterpart of shader-compiler-backend/src/lib.rs:191-257 (Compiler::run
// returning CompiledCode, CompiledCode::get by key); the pipeline compile is
// the one vkCreateGraphicsPipelines runs (translate, lower to HIP, hiprtc,
// the on-disk code-object cache).
#include "kazan/kazan_backend.h"

#include <hip/hip_runtime_api.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../engine/engine.h"

namespace kz {
bool pipeline_from_desc(const KazanPipelineDesc *d, FixedState &fx, std::string &source, std::string &err);
}

struct KazanCompiledCode {
    hipModule_t module = nullptr;
    std::vector<char> code;
    std::mutex mu;  // guards the lookup table (get may be called from several threads)
    std::map<std::string, hipFunction_t> functions;
};

namespace {

int report(const std::string &msg, char *log, size_t cap, size_t *log_len, int rc) {
    if (log_len) *log_len = msg.size();
    if (log && cap) {
        const size_t n = msg.size() < cap - 1 ? msg.size() : cap - 1;
        std::memcpy(log, msg.data(), n);
        log[n] = 0;
    }
    return rc;
}

// a gfx950 ELF code object, or a clang offload bundle holding one (what
// hipcc --genco writes)
bool is_code_object(const std::vector<char> &code) {
    static const char kBundle[] = "__CLANG_OFFLOAD_BUNDLE__";
    return kz::is_amdgpu_code_object(code.data(), code.size()) ||
           (code.size() > sizeof kBundle && std::memcmp(code.data(), kBundle, sizeof kBundle - 1) == 0);
}

int load(std::vector<char> code, KazanCompiledCode **out, char *log, size_t cap, size_t *log_len) {
    if (!is_code_object(code)) return report("not an AMDGPU code object", log, cap, log_len, -2);
    auto *c = new KazanCompiledCode;
    c->code = std::move(code);
    const hipError_t e = hipModuleLoadData(&c->module, c->code.data());
    if (e != hipSuccess) {
        delete c;
        return report(std::string("hipModuleLoadData: ") + hipGetErrorString(e), log, cap, log_len, -3);
    }
    *out = c;
    return report("", log, cap, log_len, 0);
}

}  // namespace

extern "C" const char *kazan_backend_name(void) { return "hip-gfx950"; }

extern "C" int kazan_backend_compile(const KazanPipelineDesc *desc, uint32_t optimization, KazanCompiledCode **out,
                                     char *log, size_t cap, size_t *log_len) {
    if (!desc || !out) return report("null argument", log, cap, log_len, -1);
    *out = nullptr;
    kz::FixedState fx;
    std::string source, err;
    if (!kz::pipeline_from_desc(desc, fx, source, err)) return report(err, log, cap, log_len, -1);
    fx.optimize = optimization != 0;
    std::vector<char> code;
    std::string clog;
    if (!kz::compile_pipeline_code(fx, source, code, clog, nullptr)) return report(clog, log, cap, log_len, -1);
    return load(std::move(code), out, log, cap, log_len);
}

extern "C" int kazan_code_load(const void *code_object, size_t bytes, KazanCompiledCode **out, char *log, size_t cap,
                               size_t *log_len) {
    if (!code_object || !out) return report("null argument", log, cap, log_len, -1);
    *out = nullptr;
    const char *p = (const char *)code_object;
    return load(std::vector<char>(p, p + bytes), out, log, cap, log_len);
}

extern "C" void *kazan_code_get(KazanCompiledCode *code, const char *key) {
    if (!code || !key) return nullptr;
    std::lock_guard<std::mutex> lock(code->mu);
    auto it = code->functions.find(key);
    if (it != code->functions.end()) return (void *)it->second;
    hipFunction_t f = nullptr;
    if (hipModuleGetFunction(&f, code->module, key) != hipSuccess) {
        (void)hipGetLastError();  // an absent kernel is an answer, not a sticky error
        return nullptr;
    }
    code->functions[key] = f;
    return (void *)f;
}

extern "C" void *kazan_code_module(const KazanCompiledCode *code) { return code ? (void *)code->module : nullptr; }

extern "C" const void *kazan_code_object(const KazanCompiledCode *code, size_t *bytes) {
    if (bytes) *bytes = code ? code->code.size() : 0;
    return code ? code->code.data() : nullptr;
}

extern "C" int kazan_code_launch(KazanCompiledCode *code, const char *key, const uint32_t grid[3],
                                 const uint32_t block[3], uint32_t shared_bytes, void *stream, const void *args,
                                 size_t args_bytes) {
    if (!code || !key || !grid || !block || (!args && args_bytes)) return -1;
    hipFunction_t f = (hipFunction_t)kazan_code_get(code, key);
    if (!f) return -2;
    size_t size = args_bytes;
    void *config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<void *>(args), HIP_LAUNCH_PARAM_BUFFER_SIZE, &size,
                      HIP_LAUNCH_PARAM_END};
    const hipError_t e = hipModuleLaunchKernel(f, grid[0], grid[1], grid[2], block[0], block[1], block[2],
                                               shared_bytes, (hipStream_t)stream, nullptr, args_bytes ? config : nullptr);
    return e == hipSuccess ? 0 : -3;
}

extern "C" void kazan_code_destroy(KazanCompiledCode *code) {
    if (!code) return;
    if (code->module) hipModuleUnload(code->module);
    delete code;
}
