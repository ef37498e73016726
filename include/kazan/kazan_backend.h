/*
 * kazan_backend.h — the shader-compiler backend as a C ABI: compile a
 * pipeline's shaders to a gfx950 code object, load it, look its kernels up
 * by key and launch them.
 *
 * Counterpart of the reference's backend plugin API
 * (shader-compiler-backend/src/lib.rs:191-257): Compiler::name,
 * Compiler::run(user, config) -> CompiledCode, CompiledCode::get(key) ->
 * function. The reference's CompiledCode hands back host function
 * pointers; device code is a hipModule_t and its kernels hipFunction_t
 * values, keyed here by kernel name. The pipeline kernels a compiled
 * pipeline exports are "kz_bin_count", "kz_raster" and
 * "kz_raster_persistent" (their launch arguments: one KzFrame, the engine's
 * frame record); any other kernel a loaded code object holds is found by
 * its name too.
 *
 *   kazan_backend_name          Compiler::name
 *   kazan_backend_compile       Compiler::run: SPIR-V -> IR -> HIP -> code
 *                               object (optimization: 0 = none, 1 = normal,
 *                               OptimizationMode in lib.rs), loaded on the
 *                               calling thread's current HIP device
 *   kazan_code_load             a code object's bytes (an AMDGPU ELF, or a
 *                               clang offload bundle holding one) loaded as
 *                               compiled code
 *   kazan_code_get              CompiledCode::get: the kernel for `key`
 *                               (a hipFunction_t), NULL when absent
 *   kazan_code_module           the hipModule_t
 *   kazan_code_object           the code object's bytes
 *   kazan_code_launch           launches kernel `key` (hipModuleLaunchKernel
 *                               with the argument bytes packed as one struct)
 *   kazan_code_destroy          unloads the module (CompiledCode dropped)
 *
 * Integer returns: 0 on success, negative on failure, with a message in
 * `log` when one is given (truncated to `cap`, full length in *log_len).
 */
#ifndef KAZAN_BACKEND_H
#define KAZAN_BACKEND_H

#include <stddef.h>
#include <stdint.h>

#include "kazan_tools.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct KazanCompiledCode KazanCompiledCode;

KAZAN_EXPORT const char *kazan_backend_name(void);
KAZAN_EXPORT int kazan_backend_compile(const KazanPipelineDesc *desc, uint32_t optimization,
                                       KazanCompiledCode **out, char *log, size_t cap, size_t *log_len);
KAZAN_EXPORT int kazan_code_load(const void *code_object, size_t bytes, KazanCompiledCode **out, char *log,
                                 size_t cap, size_t *log_len);
KAZAN_EXPORT void *kazan_code_get(KazanCompiledCode *code, const char *key);
KAZAN_EXPORT void *kazan_code_module(const KazanCompiledCode *code);
KAZAN_EXPORT const void *kazan_code_object(const KazanCompiledCode *code, size_t *bytes);
/* grid and block in workgroups and threads (x, y, z); `stream` is a
 * hipStream_t (NULL: the null stream); `args` are the kernel's arguments laid
 * out as one struct of `args_bytes` bytes */
KAZAN_EXPORT int kazan_code_launch(KazanCompiledCode *code, const char *key, const uint32_t grid[3],
                                   const uint32_t block[3], uint32_t shared_bytes, void *stream, const void *args,
                                   size_t args_bytes);
KAZAN_EXPORT void kazan_code_destroy(KazanCompiledCode *code);

#ifdef __cplusplus
}
#endif

#endif /* KAZAN_BACKEND_H */
