// backend_hip.h — shader IR -> HIP device code for gfx950.
//
// The reference's backend plugin API (shader-compiler-backend/src/lib.rs:
// Compiler::run at :245-257, CompiledCode::get at :200-205) hands back host
// function pointers from an LLVM 7 CPU JIT. A GPU backend cannot return
// callable host code, so this backend returns HIP C++ source: the vertex and
// fragment stages as device functions, plus the vertex-fetch and varying
// linkage glue, which pipeline creation splices into the raster kernel
// template and compiles to a gfx950 code object (engine/pipeline_compiler).
//
// Calling convention (the counterpart of the planned
// shader_main(GlobalState*, InvocationState*, InvocationInput*,
// InvocationOutput*) ABI, docs/Example Translation from SPIR-V to LLVM IR.md
// :190-198): one GPU lane is one invocation; the user input and output
// interface blocks are per-lane float arrays laid out at location*16 +
// component*4 bytes, exactly as the IR's interface blocks describe.
#pragma once

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "ir.h"

namespace kz {

struct VertexBindingDesc {
    uint32_t binding = 0, stride = 0, input_rate = 0;
};

struct VertexAttributeDesc {
    uint32_t location = 0, binding = 0, format = 0, offset = 0;
};

struct PipelineCodegenInput {
    const ir::Module *vs = nullptr;
    const ir::Module *fs = nullptr;  // may be null: no colour output
    std::vector<VertexBindingDesc> bindings;
    std::vector<VertexAttributeDesc> attributes;
    // (set, binding) -> slot in KzFrame::ubo / KzFrame::tex
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> ubo_slot, tex_slot;
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> ssbo_slot;  // compute: storage buffers
};

struct PipelineCodegenOutput {
    std::string source;       // device functions kz_vertex / kz_fragment
    uint32_t num_varyings = 0;  // packed varying floats (KZ_NV)
    bool fs_reads_frag_coord = false;
};

bool generate_pipeline_source(const PipelineCodegenInput &in, PipelineCodegenOutput &out, std::string &err);

// A compute pipeline's source: the shader as device function kz_cs_main,
// its workgroup memory, and the kernel kz_compute(KzComputeFrame) that runs
// one workgroup per dispatched group (the module's local size as its block).
// Uses in.ubo_slot and in.ssbo_slot for the descriptor bindings.
bool generate_compute_source(const ir::Module &cs, const PipelineCodegenInput &in, std::string &out,
                             std::string &err);

}  // namespace kz
