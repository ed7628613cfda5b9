// Wavefront path tracer entry points (see render.hip).
#pragma once

#include "common.hpp"

namespace render {
int set_shading(trace_ctx* ctx, trace_mesh* mesh, const double* h_normals,
                const int32_t* h_mat_index, const float* h_materials, int32_t nmat, const float* h_params);
// Recompute geometric vertex normals from the current positions (asynchronous).
int refresh_geo_normals(trace_ctx* ctx, trace_mesh* mesh);
// Enqueues a render of prm->spp samples per pixel; writes per-pixel sums,
// sums of squares and the last sample (d_last may be null), 3 floats each.
int render(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh, const trace_camera* cam,
           const trace_render_params* prm, float* d_sum, float* d_sumsq, float* d_last);
// nscene images in one wavefront; trees[c] (all views of one forest, or one
// tree) and cams[c] belong to scene c, which draws the Philox stream of
// prm->seed + c.  Outputs are scene-major.
// With d_keep (W*H*spp float4 per scene, scene-major, then sample-major),
// the render traces the adjoint's summary sweep instead, derives the same
// image from it and stores every path's summary for backward_scenes.
int render_scenes(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* const* trees,
                  int32_t nscene, const trace_camera* cams, const trace_render_params* prm,
                  float* d_sum, float* d_sumsq, float* d_last, float4* d_keep = nullptr);
// Gradient of a loss on the mean image (sum / spp) of render_scenes with the
// same arguments, with respect to the per-triangle albedo and emission
// (3 floats per triangle each, overwritten); d_grad_image = dLoss/d(mean
// image), scene-major.  Path replay: every path is traced once to find its
// end, then once more to hand each vertex its share; no path is stored.  With
// d_keep (the summaries of render_scenes with the same arguments and
// parameters) the first sweep is skipped.
int backward_scenes(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* const* trees,
                    int32_t nscene, const trace_camera* cams, const trace_render_params* prm,
                    const float* d_grad_image, float* d_grad_albedo, float* d_grad_emission,
                    const float4* d_keep = nullptr);
}  // namespace render
