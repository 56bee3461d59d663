// Host-side launchers of the device kernels (one declaration per kernel family).
#pragma once

#include "internal.hpp"

namespace cdk {

// build.hip
BuildResult build_soups(cd_ctx* ctx, const float* tri9, const uint32_t* counts, uint32_t nmesh);

// plain.hip
hipError_t launch_zero(RunCounters* c, hipStream_t st);
hipError_t launch_tritri_plain(const float* a9, const float* b9, uint64_t n, int32_t* flags, float* seg6,
                               hipStream_t st);
hipError_t launch_traverse_plain(const QueryRec* qs, uint32_t nq, const MeshRec* meshes, const NodeRec* nodes,
                                 uint4* lp, uint64_t lp_cap, RunCounters* ctr, hipStream_t st);
hipError_t launch_narrow_plain(const uint4* lps, uint64_t nlp, uint32_t leaf_max, const QueryRec* qs,
                               const MeshRec* meshes, const float* tris, cd_contact* out, uint64_t cap,
                               RunCounters* ctr, hipStream_t st);

}  // namespace cdk
