// Wavefront path tracer (ImageRenderer.RenderToBuffer, render.go:91-129, with
// DefaultClosestHitShader, render.go:209-224).
//
// Paths are processed in batches of whole samples.  Per batch:
//   camera_kernel   one camera ray per (sample, pixel), jittered as
//                   RandomSampling does (render.go:201-205)
//   bounce_kernel   for segment k = 0 .. max_depth+1: closest hit of every
//                   live path, then shade -- miss adds throughput x miss
//                   colour; a hit at vertex k > max_depth ends the path with
//                   nothing added; a light adds throughput x emission and ends
//                   the path; a diffuse hit scatters (material.go:29-38) and
//                   the path is appended to the next ray queue (wave-aggregated
//                   atomic compaction)
//   accumulate      per pixel: sum, sum of squares and last sample, in sample
//                   order (deterministic)
// Forward accumulation of throughput x emission gives the same value as the
// reference's recursion L = Le + albedo (x) L_next.
#include <algorithm>
#include <vector>

#include "philox.hpp"
#include "render.hpp"
#include "scan.hpp"
#include "traverse_dev.hpp"

namespace render {


struct Cam {
  float o[3], ll[3], h[3], v[3];
};

__global__ void geo_normals_kernel(const double* __restrict__ pos, int64_t n, float* __restrict__ nrm) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const double* p = pos + 9 * i;
  if (prim_is_sphere(p)) {  // a sphere's shading normal comes from its centre
    for (int c = 0; c < 9; ++c) nrm[9 * i + c] = 0.0f;
    return;
  }
  // calcNormal(point, right, left) = unit((right - point) x (left - point)),
  // primitive.go:116-138, for each vertex in turn.
  for (int c = 0; c < 3; ++c) {
    const double* a = p + 3 * c;
    const double* r = p + 3 * ((c + 1) % 3);
    const double* l = p + 3 * ((c + 2) % 3);
    const double bx = r[0] - a[0], by = r[1] - a[1], bz = r[2] - a[2];
    const double ax = l[0] - a[0], ay = l[1] - a[1], az = l[2] - a[2];
    const double cx = by * az - bz * ay, cy = bz * ax - bx * az, cz = bx * ay - by * ax;
    const double len = sqrt(cx * cx + cy * cy + cz * cz);
    nrm[9 * i + 3 * c + 0] = float(cx / len);
    nrm[9 * i + 3 * c + 1] = float(cy / len);
    nrm[9 * i + 3 * c + 2] = float(cz / len);
  }
}

__global__ void d2f_kernel(const double* __restrict__ in, size_t n, float* __restrict__ out) {
  const size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = float(in[i]);
}

// One image of a (possibly batched) render: its camera, its Philox key and
// the header of the tree it is traced against.  Paths of a batch are ordered
// scene-major: path p belongs to scene p / pps (pps = paths per scene).
struct Scene {
  Cam cam;
  uint64_t seed;
  const BvhHeader* hdr;
};

__global__ void scene_fill_kernel(Scene s, Scene* out) { *out = s; }

// Per-triangle parameters and the adjoint's inputs and outputs.  mode 0 is the
// plain render; mode 1 (summary) traces the same paths and leaves, per path,
// S = E (x) prod(non-zero albedos) and flag bits in rad; mode 2 (replay)
// traces them once more and adds each vertex's share of the gradient.
struct Adj {
  const float* alb;    // Diffuse albedo per triangle (3 floats), or null: material colour
  const float* emi;    // Light emission per triangle (3 floats), or null
  const float* g_img;  // replay: dLoss/d(mean image), scene-major, 3 per pixel
  float* g_alb;        // replay: gradient outputs, 3 per triangle
  float* g_emi;
  float inv_spp;
  int32_t mode;
  int64_t wh;
};
constexpr int32_t kModePrimal = 0, kModeSummary = 1, kModeReplay = 2;
// summary flag bits (rad.w / thr.w hold them as raw bits): per channel, the
// number of zero albedos on the path so far (2 bits, saturating at 2)
constexpr uint32_t kFlagLight = 64u;    // the path ended on a light
constexpr uint32_t kFlagRadiant = 128u; // the radiance E where it ended is non-zero

__device__ inline float3 tri_colour(const Adj& a, float4 m, int32_t prim) {
  const float* o = m.x == float(TRACE_MAT_LIGHT) ? a.emi : a.alb;
  if (o) return make_float3(o[3 * size_t(prim)], o[3 * size_t(prim) + 1], o[3 * size_t(prim) + 2]);
  return make_float3(m.y, m.z, m.w);
}

// out[3 idx + c] += v_c over the active lanes of a wave.  Lanes that share an
// index (a wall, a light seen by a whole tile) are summed in registers first:
// up to four distinct indices are reduced that way, the rest go straight to
// memory.
__device__ inline void wave_add3(float* out, int32_t idx, float x, float y, float z, bool active) {
  const int lane = __lane_id();
  uint64_t todo = ballot64(active);
  for (int it = 0; it < 4 && todo; ++it) {
    const int leader = __ffsll((unsigned long long)todo) - 1;
    const int32_t key = __shfl(idx, leader, 64);
    const bool in = ((todo >> lane) & 1ull) && idx == key;
    const uint64_t grp = ballot64(in);
    float sx = in ? x : 0.0f, sy = in ? y : 0.0f, sz = in ? z : 0.0f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      sx += __shfl_xor(sx, o, 64);
      sy += __shfl_xor(sy, o, 64);
      sz += __shfl_xor(sz, o, 64);
    }
    if (lane == leader) {
      float* d = out + 3 * size_t(key);
      if (sx != 0.0f) unsafeAtomicAdd(d, sx);
      if (sy != 0.0f) unsafeAtomicAdd(d + 1, sy);
      if (sz != 0.0f) unsafeAtomicAdd(d + 2, sz);
    }
    todo &= ~grp;
  }
  if ((todo >> lane) & 1ull) {
    float* d = out + 3 * size_t(idx);
    if (x != 0.0f) unsafeAtomicAdd(d, x);
    if (y != 0.0f) unsafeAtomicAdd(d + 1, y);
    if (z != 0.0f) unsafeAtomicAdd(d + 2, z);
  }
}

// Camera ray of path q of a scene (camera.go:95-97, RandomSampling render.go:201-205).
__device__ inline void camera_ray(const Scene& sc, int32_t W, int32_t H, int64_t q, int64_t path_base,
                                  float4& o, float4& d) {
  const Cam& cam = sc.cam;
  const int64_t wh = int64_t(W) * H;
  const int64_t pix = q % wh;
  const int32_t x = int32_t(pix % W), y = int32_t(pix / W);
  const Philox4 r = path_draw(sc.seed, uint64_t(path_base + q), 0u);
  const float s = (float(x) + u01(r.v[0])) / float(W - 1);
  const float t = (float(y) + u01(r.v[1])) / float(H - 1);
  const float dx = cam.ll[0] + s * cam.h[0] + t * cam.v[0] - cam.o[0];
  const float dy = cam.ll[1] + s * cam.h[1] + t * cam.v[1] - cam.o[1];
  const float dz = cam.ll[2] + s * cam.h[2] + t * cam.v[2] - cam.o[2];
  o = make_float4(cam.o[0], cam.o[1], cam.o[2], 0.0f);
  d = make_float4(dx, dy, dz, 0.0f);
}

__global__ __launch_bounds__(256) void camera_kernel(const Scene* __restrict__ scenes, int32_t W,
                                                     int32_t H, int64_t n_paths, int64_t pps,
                                                     int64_t path_base, float4* __restrict__ ro,
                                                     float4* __restrict__ rd,
                                                     float4* __restrict__ thr,
                                                     float4* __restrict__ rad,
                                                     int32_t* __restrict__ queue,
                                                     uint32_t* __restrict__ counts) {
  const int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (p == 0) {
    counts[0] = uint32_t(n_paths);
    counts[1] = 0;
  }
  if (p >= n_paths) return;
  float4 o, d;
  camera_ray(scenes[p / pps], W, H, p % pps, path_base, o, d);
  ro[p] = o;
  rd[p] = d;
  thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
  rad[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  queue[p] = int32_t(p);
}

// Replay start: re-creates the camera ray of every path whose summary can
// carry a gradient (it ended on a light, or with non-zero radiance, and its
// pixel's image gradient is non-zero) and queues it; rad keeps the summary.
// With `keep`, the summaries come from a primal that kept them (at
// keep[scene * keep_scene + keep_off + q]) and are copied into rad for the
// replayed paths.
__global__ __launch_bounds__(256) void replay_camera_kernel(
    const Scene* __restrict__ scenes, int32_t W, int32_t H, int64_t n_paths, int64_t pps,
    int64_t path_base, const float* __restrict__ g_img, float4* __restrict__ ro,
    float4* __restrict__ rd, float4* __restrict__ thr, float4* __restrict__ rad,
    int32_t* __restrict__ queue, uint32_t* __restrict__ count, const float4* __restrict__ keep,
    int64_t keep_scene, int64_t keep_off) {
  const int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  bool push = false;
  if (p < n_paths) {
    const int64_t wh = int64_t(W) * H;
    const int64_t scene = p / pps, q = p % pps;
    const float* g = g_img + 3 * (scene * wh + q % wh);
    const float4 S = keep ? keep[scene * keep_scene + keep_off + q] : rad[p];
    const uint32_t flags = __float_as_uint(S.w);
    push = (flags & (kFlagLight | kFlagRadiant)) && (g[0] != 0.0f || g[1] != 0.0f || g[2] != 0.0f);
    if (push) {
      if (keep) rad[p] = S;
      float4 o, d;
      camera_ray(scenes[scene], W, H, q, path_base, o, d);
      ro[p] = o;
      rd[p] = d;
      thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    }
  }
  const uint64_t b = ballot64(push);
  if (b) {
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)b) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, uint32_t(__popcll(b)));
    base = __shfl(base, leader, 64);
    if (push) queue[base + __popcll(b & lanemask_lt())] = int32_t(p);
  }
}

// Moves a surface point a few float ULPs along n (Waechter & Binder, "A Fast
// and Robust Method for Avoiding Self-Intersection", Ray Tracing Gems, 2019):
// an integer offset of the float bits away from zero, and a small absolute
// offset near the origin where ULPs vanish.  The reference works in FP64,
// where a new ray starting on a triangle never re-enters it; in FP32 the
// rounded hit point can sit on the wrong side of the plane, and a grazing
// scattered ray would then hit the surface it left (on a closed mesh, from
// inside).  Offsetting to the side the new ray leaves toward restores the
// FP64 behaviour.
__device__ inline float offset_coord(float p, float n) {
  constexpr float kOrigin = 1.0f / 32.0f, kFloatScale = 1.0f / 65536.0f, kIntScale = 32.0f;
  const int oi = int(kIntScale * n);
  const float pi = __int_as_float(__float_as_int(p) + (p < 0.0f ? -oi : oi));
  return fabsf(p) < kOrigin ? p + kFloatScale * n : pi;
}

// render.go:228-246
__device__ inline float3 miss_color(int miss, float dx, float dy, float dz) {
  switch (miss) {
    case TRACE_MISS_WHITE:
      return make_float3(1.0f, 1.0f, 1.0f);
    case TRACE_MISS_SKY: {
      const float len = sqrtf(dx * dx + dy * dy + dz * dz);
      const float t = 0.5f * (dy / len + 1.0f);
      return make_float3(0.8f * (1.0f - t) + 0.25f * t, 0.8f * (1.0f - t) + 0.35f * t,
                         0.8f * (1.0f - t) + 0.5f * t);
    }
    case TRACE_MISS_SUN:
      return make_float3(1.0f, 0.96f, 0.78f);
    default:
      return make_float3(0.0f, 0.0f, 0.0f);
  }
}

// Everything the kernels below share about one path segment.
struct SegArgs {
  const float* tris;
  const int32_t* perm;
  const Scene* scenes;
  int64_t pps;
  const float* nrm;
  const int32_t* mat_idx;
  const float4* mats;
  const float* mat_param;  // per material: Reflective.Diffusion / Refractive.Ratio
  int32_t k, max_depth, miss;
  int64_t path_base;
  float4 *ro, *rd, *thr, *rad;
  unsigned long long* seg_total;  // segments traced since the call began
  unsigned long long* visits;     // counting mode: wide-node visits, triangle tests (else null)
};

// Shades segment k of path p after its closest hit h (primitive.go:145-150,
// render.go:209-246, material.go:13-38): miss, light or diffuse scatter, in
// the mode adj names.  Sets push for a path that continues, and in replay
// mode the triangles that receive a gradient (al_tri, em_tri) with weights.
template <bool SPH, bool SPEC>
__device__ inline void shade_segment(const SegArgs& g, const Adj& adj, int32_t p, float4 o4, float4 d4,
                                     const traverse::Hit& h, bool& push, int32_t& al_tri, float4& al_w,
                                     int32_t& em_tri, float4& em_w) {
  const float* __restrict__ tris = g.tris;
  const int32_t* __restrict__ perm = g.perm;
  const float* __restrict__ nrm = g.nrm;
  const int32_t* __restrict__ mat_idx = g.mat_idx;
  const float4* __restrict__ mats = g.mats;
  float4* __restrict__ ro = g.ro;
  float4* __restrict__ rd = g.rd;
  float4* __restrict__ thr = g.thr;
  float4* __restrict__ rad = g.rad;
  const int32_t k = g.k, max_depth = g.max_depth, miss = g.miss;
  const Scene& sc = g.scenes[p / g.pps];
  const uint64_t seed = sc.seed;
  const int64_t path_id = g.path_base + p % g.pps;
  const float4 T = thr[p];
  if (h.pos < 0) {
    const float3 m = miss_color(miss, d4.x, d4.y, d4.z);
    if (adj.mode == kModePrimal) {
      float4 L = rad[p];
      L.x += T.x * m.x;
      L.y += T.y * m.y;
      L.z += T.z * m.z;
      rad[p] = L;
    } else if (adj.mode == kModeSummary) {
      const uint32_t f = __float_as_uint(T.w) | ((m.x != 0.0f || m.y != 0.0f || m.z != 0.0f) ? kFlagRadiant : 0u);
      rad[p] = make_float4(m.x * T.x, m.y * T.y, m.z * T.z, __uint_as_float(f));
    }
  } else if (k <= max_depth) {
    const int32_t prim = perm[h.pos];
    const float4 m = mats[mat_idx[prim]];
    const float3 c = tri_colour(adj, m, prim);
    if (m.x == float(TRACE_MAT_LIGHT)) {
      if (adj.mode == kModePrimal) {
        float4 L = rad[p];
        L.x += T.x * c.x;
        L.y += T.y * c.y;
        L.z += T.z * c.z;
        rad[p] = L;
      } else if (adj.mode == kModeSummary) {
        const uint32_t f = __float_as_uint(T.w) | kFlagLight |
                           ((c.x != 0.0f || c.y != 0.0f || c.z != 0.0f) ? kFlagRadiant : 0u);
        rad[p] = make_float4(c.x * T.x, c.y * T.y, c.z * T.z, __uint_as_float(f));
      } else {
        // dL/dEmission = product of the albedos before the light
        em_tri = prim;
        em_w = T;
      }
    } else {
      const float* tv = tris + 9 * size_t(h.pos);
      const bool sph = SPH && prim_is_sphere(tv);
      // hit point = ray.position(t) (camera.go:108-111), evaluated in FP64
      const float hx = float(double(o4.x) + double(d4.x) * h.t64);
      const float hy = float(double(o4.y) + double(d4.y) * h.t64);
      const float hz = float(double(o4.z) + double(d4.z) * h.t64);
      float nx, ny, nz;
      if (sph) {
        // Sphere: (point - centre) * (1 / radius) (primitive.go:80-84)
        const float inv = 1.0f / tv[3];
        nx = (hx - tv[0]) * inv;
        ny = (hy - tv[1]) * inv;
        nz = (hz - tv[2]) * inv;
      } else {
        // interpolated vertex normal (primitive.go:145-150, 200-205)
        const float* n = nrm + 9 * size_t(prim);
        const float w0 = 1.0f - h.u - h.v;
        nx = (n[3] * h.u + n[6] * h.v) + n[0] * w0;
        ny = (n[4] * h.u + n[7] * h.v) + n[1] * w0;
        nz = (n[5] * h.u + n[8] * h.v) + n[2] * w0;
      }
      if (!h.front) {  // flipped to face the ray
        nx = -nx;
        ny = -ny;
        nz = -nz;
      }
      // a uniform unit vector (util.go:15-26 draws a uniform direction)
      const Philox4 r = path_draw(seed, uint64_t(path_id), uint32_t(k + 1));
      const float z = 1.0f - 2.0f * u01(r.v[0]);
      const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
      float sphi, cphi;
      sincosf(6.2831853071795864f * u01(r.v[1]), &sphi, &cphi);
      const float ux = rr * cphi, uy = rr * sphi, uz = z;
      const int kind = SPEC ? int(m.x) : TRACE_MAT_DIFFUSE;
      float sx, sy, sz;
      if (kind == TRACE_MAT_DIFFUSE) {
        // Diffuse.scatter (material.go:29-38): n + unit vector, n if that is ~0
        sx = nx + ux;
        sy = ny + uy;
        sz = nz + uz;
        if (fabsf(sx) <= 1e-7f && fabsf(sy) <= 1e-7f && fabsf(sz) <= 1e-7f) {
          sx = nx;
          sy = ny;
          sz = nz;
        }
      } else {
        // unit incoming direction (math.go Unit)
        const float dl = sqrtf(d4.x * d4.x + d4.y * d4.y + d4.z * d4.z);
        const float vx = d4.x / dl, vy = d4.y / dl, vz = d4.z / dl;
        const float vn = vx * nx + vy * ny + vz * nz;
        // reflect (material.go:90-92): v - n (2 v.n)
        const float rx = vx - nx * (vn * 2.0f), ry = vy - ny * (vn * 2.0f), rz = vz - nz * (vn * 2.0f);
        const float prm = g.mat_param[mat_idx[prim]];
        if (kind == TRACE_MAT_REFLECTIVE) {
          // Reflective.scatter (material.go:50-54): reflection + Diffusion x unit
          // vector; no scatter (the path ends dark) unless reflected . n > 0
          if (!(rx * nx + ry * ny + rz * nz > 0.0f)) return;
          sx = rx + ux * prm;
          sy = ry + uy * prm;
          sz = rz + uz * prm;
        } else {
          // Refractive.scatter (material.go:65-84) with Schlick's reflectance
          // (material.go:101-106); the draw deciding reflection is this
          // vertex's third Philox word
          const float ratio = h.front ? 1.0f / prm : prm;
          const float cos_t = fminf(-vx * nx + -vy * ny + -vz * nz, 1.0f);
          const float sin_t = sqrtf(1.0f - cos_t * cos_t);
          float r0 = (1.0f - ratio) / (1.0f + ratio);
          r0 = r0 * r0;
          const float om = 1.0f - cos_t;
          const float refl = r0 + (1.0f - r0) * (om * om * om * om * om);
          if (ratio * sin_t > 1.0f || refl > u01(r.v[2])) {
            sx = rx;
            sy = ry;
            sz = rz;
          } else {
            // refract (material.go:94-99)
            const float px = (vx + nx * cos_t) * ratio, py = (vy + ny * cos_t) * ratio,
                        pz = (vz + nz * cos_t) * ratio;
            const float par = -sqrtf(fabsf(1.0f - (px * px + py * py + pz * pz)));
            sx = px + nx * par;
            sy = py + ny * par;
            sz = pz + nz * par;
          }
        }
      }
      // the new origin is the hit point, kept on the side of the surface
      // (triangle plane or sphere tangent plane) the new ray leaves to
      float gx, gy, gz;
      if (sph) {
        gx = hx - tv[0];
        gy = hy - tv[1];
        gz = hz - tv[2];
      } else {
        const float e1x = tv[3] - tv[0], e1y = tv[4] - tv[1], e1z = tv[5] - tv[2];
        const float e2x = tv[6] - tv[0], e2y = tv[7] - tv[1], e2z = tv[8] - tv[2];
        gx = e1y * e2z - e1z * e2y;
        gy = e1z * e2x - e1x * e2z;
        gz = e1x * e2y - e1y * e2x;
      }
      const float gl = sqrtf(gx * gx + gy * gy + gz * gz);
      const float gs = (gx * sx + gy * sy + gz * sz) >= 0.0f ? 1.0f / gl : -1.0f / gl;
      gx *= gs;
      gy *= gs;
      gz *= gs;
      ro[p] = make_float4(offset_coord(hx, gx), offset_coord(hy, gy), offset_coord(hz, gz), 0.0f);
      rd[p] = make_float4(sx, sy, sz, 0.0f);
      if (adj.mode == kModeSummary) {
        // product of the non-zero albedos, and a count of the zero ones
        uint32_t f = __float_as_uint(T.w);
        float tc[3] = {T.x, T.y, T.z};
        const float cc[3] = {c.x, c.y, c.z};
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          if (cc[ch] != 0.0f) {
            tc[ch] *= cc[ch];
          } else {
            const uint32_t z = (f >> (2 * ch)) & 3u;
            if (z < 2u) f += 1u << (2 * ch);
          }
        }
        thr[p] = make_float4(tc[0], tc[1], tc[2], __uint_as_float(f));
      } else {
        thr[p] = make_float4(T.x * c.x, T.y * c.y, T.z * c.z, 0.0f);
      }
      if (adj.mode == kModeReplay && kind == TRACE_MAT_DIFFUSE) {
        // dL/da_j = S / a_j when the path has no zero albedo; with exactly
        // one zero albedo only that vertex has a gradient, S itself
        const float4 S = rad[p];
        const uint32_t f = __float_as_uint(S.w);
        const float sc3[3] = {S.x, S.y, S.z};
        const float cc[3] = {c.x, c.y, c.z};
        float gv[3];
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          const uint32_t z = (f >> (2 * ch)) & 3u;
          gv[ch] = cc[ch] != 0.0f ? (z == 0u ? sc3[ch] / cc[ch] : 0.0f) : (z == 1u ? sc3[ch] : 0.0f);
        }
        al_tri = prim;
        al_w = make_float4(gv[0], gv[1], gv[2], 0.0f);
      }
      push = true;
    }
  }
}

// Replay-mode gradient scatter and the wave-aggregated append of continuing
// paths to the next queue; every lane of the wave must take part.
__device__ inline void segment_epilogue(const Adj& adj, int64_t pps, int32_t p, bool push, int32_t al_tri,
                                        float4 al_w, int32_t em_tri, float4 em_w, int32_t* __restrict__ q_out,
                                        uint32_t* __restrict__ count_out) {
  if (adj.mode == kModeReplay) {
    // weight of the path: dLoss/d(mean image) at its pixel / spp
    float wx = 0.0f, wy = 0.0f, wz = 0.0f;
    if (al_tri >= 0 || em_tri >= 0) {
      const int64_t q = p % pps;
      const float* g = adj.g_img + 3 * ((p / pps) * adj.wh + q % adj.wh);
      wx = g[0] * adj.inv_spp;
      wy = g[1] * adj.inv_spp;
      wz = g[2] * adj.inv_spp;
    }
    wave_add3(adj.g_alb, al_tri, wx * al_w.x, wy * al_w.y, wz * al_w.z, al_tri >= 0);
    wave_add3(adj.g_emi, em_tri, wx * em_w.x, wy * em_w.y, wz * em_w.z, em_tri >= 0);
  }
  // wave-aggregated append to the next queue
  const uint64_t b = ballot64(push);
  if (b) {
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)b) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count_out, uint32_t(__popcll(b)));
    base = __shfl(base, leader, 64);
    if (push) q_out[base + __popcll(b & lanemask_lt())] = p;
  }
}

// One thread per (scene, pixel): the batch's samples in sample order.  Output
// is scene-major: scene c's image starts at 3 * c * wh.
//
// With `keep`, rad holds path summaries (the summary sweep) instead of
// radiance: channel c of a path's radiance is its summary's when no albedo on
// the path was zero in that channel (the summary multiplies the non-zero
// albedos in the primal's order, so the value is the primal's bit for bit),
// and 0 otherwise.  The summaries themselves are stored at keep[scene *
// keep_scene + keep_off + q] for the backward pass.
__global__ void accumulate_kernel(const float4* __restrict__ rad, int64_t wh, int32_t nscene,
                                  int32_t nsamp, int32_t first_batch, float* __restrict__ sum,
                                  float* __restrict__ sumsq, float* __restrict__ last,
                                  float4* __restrict__ keep, int64_t keep_scene, int64_t keep_off) {
  const int64_t qx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (qx >= wh * nscene) return;
  const int64_t scene = qx / wh, pix = qx % wh;
  const float4* base = rad + scene * nsamp * wh + pix;
  const int64_t px = qx;  // output index (scene-major)
  float s0 = 0, s1 = 0, s2 = 0, q0 = 0, q1 = 0, q2 = 0;
  float4 v = make_float4(0, 0, 0, 0);
  for (int32_t s = 0; s < nsamp; ++s) {
    v = base[int64_t(s) * wh];
    if (keep) {
      keep[scene * keep_scene + keep_off + int64_t(s) * wh + pix] = v;
      const uint32_t f = __float_as_uint(v.w);
      v.x = (f & 3u) == 0u ? v.x : 0.0f;
      v.y = ((f >> 2) & 3u) == 0u ? v.y : 0.0f;
      v.z = ((f >> 4) & 3u) == 0u ? v.z : 0.0f;
    }
    s0 += v.x;
    s1 += v.y;
    s2 += v.z;
    q0 += v.x * v.x;
    q1 += v.y * v.y;
    q2 += v.z * v.z;
  }
  float* S = sum + 3 * px;
  float* Q = sumsq + 3 * px;
  if (first_batch) {
    S[0] = s0; S[1] = s1; S[2] = s2;
    Q[0] = q0; Q[1] = q1; Q[2] = q2;
  } else {
    S[0] += s0; S[1] += s1; S[2] += s2;
    Q[0] += q0; Q[1] += q1; Q[2] += q2;
  }
  if (last) {
    last[3 * px + 0] = v.x;
    last[3 * px + 1] = v.y;
    last[3 * px + 2] = v.z;
  }
}

int set_shading(trace_ctx* ctx, trace_mesh* mesh, const double* h_normals,
                const int32_t* h_mat_index, const float* h_materials, int32_t nmat, const float* h_params) {
  const int64_t n = mesh->ntri;
  hipStream_t st = ctx->stream;
  HIP_TRY(ctx, mesh->nrm.alloc(9 * n));
  HIP_TRY(ctx, mesh->mat_idx.alloc(n));
  HIP_TRY(ctx, mesh->mats.alloc(nmat));
  HIP_TRY(ctx, mesh->mat_param.alloc(nmat));
  std::vector<float> params(h_params ? h_params : nullptr, h_params ? h_params + nmat : nullptr);
  params.resize(size_t(nmat), 0.0f);
  mesh->has_specular = false;
  for (int32_t m = 0; m < nmat; ++m)
    if (h_materials[4 * m] == float(TRACE_MAT_REFLECTIVE) || h_materials[4 * m] == float(TRACE_MAT_REFRACTIVE))
      mesh->has_specular = true;
  HIP_TRY(ctx, hipMemcpyAsync(mesh->mat_param.p, params.data(), nmat * sizeof(float), hipMemcpyHostToDevice, st));
  if (h_normals) {
    DevArray<double> tmp;
    HIP_TRY(ctx, tmp.alloc(9 * n));
    HIP_TRY(ctx, hipMemcpyAsync(tmp.p, h_normals, 9 * n * sizeof(double), hipMemcpyHostToDevice, st));
    d2f_kernel<<<int(ceil_div(9 * n, 256)), 256, 0, st>>>(tmp.p, size_t(9 * n), mesh->nrm.p);
    HIP_TRY(ctx, hipStreamSynchronize(st));
  } else {
    geo_normals_kernel<<<int(ceil_div(n, 256)), 256, 0, st>>>(mesh->pos.p, n, mesh->nrm.p);
  }
  HIP_TRY(ctx, hipMemcpyAsync(mesh->mat_idx.p, h_mat_index, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(ctx, hipMemcpyAsync(mesh->mats.p, h_materials, nmat * sizeof(float4), hipMemcpyHostToDevice, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  mesh->nmat = nmat;
  mesh->has_shading = true;
  mesh->geo_normals = h_normals == nullptr;
  return TRACE_OK;
}

int refresh_geo_normals(trace_ctx* ctx, trace_mesh* mesh) {
  const int64_t n = mesh->ntri;
  geo_normals_kernel<<<int(ceil_div(n, 256)), 256, 0, ctx->stream>>>(mesh->pos.p, n, mesh->nrm.p);
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

// Eight waves per SIMD (64 VGPRs, a few spilled): the render's late bounces
// are latency-bound, and the extra waves hide more of it than the spills cost
// (measured on the 10.4M-triangle frame: 125 ms at 6 waves, 118 at 7, 113 at 8).
// CNT: the counting form, which also sums the wide-node visits and triangle
// tests of its segments (for the byte accounting of the render's roofline).
template <int THREADS, int STACK, bool SPH, bool SPEC, bool CNT = false>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(8, 8))) void bounce_kernel(
    const WideNode* __restrict__ wnodes, const SegArgs g,
    const int32_t* __restrict__ q_in, int32_t* __restrict__ q_out, const uint32_t* __restrict__ count_in,
    uint32_t* __restrict__ count_out, uint32_t* __restrict__ d_err, const Adj adj) {
  __shared__ int32_t s_stack[traverse::kLdsStack * THREADS];
  const uint32_t count = *count_in;
  const uint32_t i = blockIdx.x * uint32_t(THREADS) + threadIdx.x;
  if (i == 0) atomicAdd(g.seg_total, static_cast<unsigned long long>(count));
  bool push = false;
  int32_t p = 0;
  int32_t al_tri = -1, em_tri = -1;  // replay: triangles that receive a gradient
  float4 al_w = make_float4(0.0f, 0.0f, 0.0f, 0.0f), em_w = al_w;
  int32_t nn = 0, nt = 0;
  if (i < count) {
    p = q_in[i];
    const BvhHeader hdr = *g.scenes[p / g.pps].hdr;
    const float4 o4 = g.ro[p];
    const float4 d4 = g.rd[p];
    traverse::Ray ray{o4.x, o4.y, o4.z, g.k == 0 ? 0.001f : 0.0001f, d4.x, d4.y, d4.z, INFINITY};
    bool overflow = false, bad = false;
    traverse::Hit h = traverse::closest_hit_wide<STACK, CNT, SPH>(
        wnodes, g.tris, hdr, ray, nn, nt, overflow, bad, s_stack + threadIdx.x, THREADS);
    if (overflow) atomicOr(d_err, TRACE_ERR_STACK);
    if (bad) atomicOr(d_err, TRACE_ERR_TREE);
    traverse::refine_hit_fp64(g.tris, ray, h);
    shade_segment<SPH, SPEC>(g, adj, p, o4, d4, h, push, al_tri, al_w, em_tri, em_w);
  }
  if constexpr (CNT) {
    for (int o = 32; o > 0; o >>= 1) {
      nn += __shfl_xor(nn, o, 64);
      nt += __shfl_xor(nt, o, 64);
    }
    if (__lane_id() == 0 && (nn | nt)) {
      atomicAdd(g.visits, static_cast<unsigned long long>(nn));
      atomicAdd(g.visits + 1, static_cast<unsigned long long>(nt));
    }
  }
  segment_epilogue(adj, g.pps, p, push, al_tri, al_w, em_tri, em_w, q_out, count_out);
}

// Split form of a segment (the autotuner's alternative to bounce_kernel): a
// persistent closest-hit pass over the queue (the scheme of
// closest_hit_persistent: waves claim chunks of queue slots from a counter and
// refill lanes as rays finish) writes one 16-byte hit record per slot, then
// shade_kernel shades slot i in thread i.  The traversal lanes never wait on
// shading; the records cost 32 bytes of traffic per segment.
constexpr uint32_t kSegChunk = 256;
// Queue order (see qcell_count_kernel): 64 origin cells, tiles of 4096 slots,
// by default for batches of at least kQOrderMin paths.
constexpr int64_t kQOrderMin = int64_t(1) << 22;
constexpr int kQCellBits = 2;
constexpr int kQCells = 1 << (3 * kQCellBits);
constexpr int kQThreads = 256;
constexpr int kQWaves = kQThreads / 64;
constexpr int kQPer = 16;
constexpr int kQTile = kQThreads * kQPer;
constexpr int kSegRefillAt = 16;

// Hit record: triangle position (-1: miss), t with the sign of the facing
// (t > 0 for every hit, + for front), FP32 barycentrics.
__device__ inline int4 pack_hit(const traverse::Hit& h) {
  return make_int4(h.pos, __float_as_int(h.front ? h.t : -h.t), __float_as_int(h.u), __float_as_int(h.v));
}

template <int THREADS, int STACK, bool SPH>
__global__ __launch_bounds__(THREADS) void trace_persistent(
    const WideNode* __restrict__ wnodes, const SegArgs g, const int32_t* __restrict__ q_in,
    const uint32_t* __restrict__ count_in, int4* __restrict__ hits, uint32_t* __restrict__ d_err,
    uint32_t* __restrict__ claim) {
  __shared__ int32_t s_stack[traverse::kLdsStack * THREADS];
  const int64_t count = *count_in;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(g.seg_total, static_cast<unsigned long long>(count));
  const int lane = __lane_id();
  const uint64_t lt = lanemask_lt();
  const float tmin = g.k == 0 ? 0.001f : 0.0001f;
  traverse::TravSpill<STACK> spill;
  traverse::TravStack<STACK> stack;
  stack.lds = traverse::lds_column(s_stack + threadIdx.x);
  stack.stride = THREADS;
  stack.spill = spill.v;
  traverse::WideState s;
  int64_t slot = 0;
  uint32_t n_bound = 0;
  int32_t n_prims = 0;
  bool active = false, pending = false, overflow = false, bad = false;
  bool fetching = true;                  // the same in every lane of the wave
  int64_t pool_next = 0, pool_end = 0;   // this wave's unclaimed queue slots
  int32_t nn = 0, nt = 0;
  while (true) {
    const uint64_t need = ballot64(!active);
    if (__popcll(need) >= kSegRefillAt || need == ~0ull) {
      if (pending) {
        hits[slot] = pack_hit(s.h);
        pending = false;
      }
      if (pool_next >= pool_end && fetching) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(claim, kSegChunk);
        base = __shfl(base, 0, 64);
        if (int64_t(base) >= count) {
          fetching = false;
        } else {
          pool_next = base;
          pool_end = base + kSegChunk < uint64_t(count) ? int64_t(base) + kSegChunk : count;
        }
      }
      const int64_t avail = pool_end - pool_next;
      if (avail > 0) {
        const int64_t rank = __popcll(need & lt);
        if (!active && rank < avail) {
          slot = pool_next + rank;
          const int32_t p = q_in[slot];
          const BvhHeader hdr = *g.scenes[p / g.pps].hdr;
          n_bound = uint32_t(hdr.n_wide);
          n_prims = hdr.n_prims;
          const float4 o4 = g.ro[p];
          const float4 d4 = g.rd[p];
          const traverse::Ray ray{o4.x, o4.y, o4.z, tmin, d4.x, d4.y, d4.z, INFINITY};
          active = traverse::wide_begin<false, SPH>(g.tris, hdr, ray, s, nt, bad);
          pending = !active;
        }
        pool_next += __popcll(need) < avail ? int64_t(__popcll(need)) : avail;
      }
    }
    if (!ballot64(active)) {
      if (!fetching && pool_next >= pool_end && !ballot64(pending)) break;
      continue;
    }
    if (active) {
      active = traverse::wide_step<STACK, false, SPH>(wnodes, g.tris, n_bound, n_prims, s, stack, nn, nt,
                                                 overflow, bad);
      pending = !active;
    }
  }
  if (overflow) atomicOr(d_err, TRACE_ERR_STACK);
  if (bad) atomicOr(d_err, TRACE_ERR_TREE);
}

template <int THREADS, bool SPH, bool SPEC>
__global__ __launch_bounds__(THREADS) void shade_kernel(const SegArgs g, const int32_t* __restrict__ q_in,
                                                        int32_t* __restrict__ q_out,
                                                        const uint32_t* __restrict__ count_in,
                                                        uint32_t* __restrict__ count_out,
                                                        const int4* __restrict__ hits, const Adj adj) {
  const uint32_t count = *count_in;
  const uint32_t i = blockIdx.x * uint32_t(THREADS) + threadIdx.x;
  bool push = false;
  int32_t p = 0;
  int32_t al_tri = -1, em_tri = -1;
  float4 al_w = make_float4(0.0f, 0.0f, 0.0f, 0.0f), em_w = al_w;
  if (i < count) {
    p = q_in[i];
    const float4 o4 = g.ro[p];
    const float4 d4 = g.rd[p];
    const int4 r = hits[i];
    const float ts = __int_as_float(r.y);
    traverse::Hit h;
    h.pos = r.x;
    h.t = fabsf(ts);
    h.t64 = h.t;
    h.front = ts > 0.0f;
    h.u = __int_as_float(r.z);
    h.v = __int_as_float(r.w);
    const traverse::Ray ray{o4.x, o4.y, o4.z, g.k == 0 ? 0.001f : 0.0001f, d4.x, d4.y, d4.z, INFINITY};
    traverse::refine_hit_fp64(g.tris, ray, h);
    shade_segment<SPH, SPEC>(g, adj, p, o4, d4, h, push, al_tri, al_w, em_tri, em_w);
  }
  segment_epilogue(adj, g.pps, p, push, al_tri, al_w, em_tri, em_w, q_out, count_out);
}

template <int THREADS, bool SPH, bool SPEC>
static void launch_segments(trace_ctx* ctx, const trace_bvh* bvh, int grid, const SegArgs& g, int32_t* qin,
                            int32_t* qout, uint32_t* cin, uint32_t* cout, int4* hits, const Adj& adj) {
  hipStream_t st = ctx->stream;
  const bool deep = bvh->depth_bound() > 63;
  if constexpr (THREADS == 64) {
    if (g.visits) {  // counting mode (64-thread workgroups only)
      if (deep)
        bounce_kernel<64, 512, SPH, SPEC, true><<<grid, 64, 0, st>>>(bvh->wnodes_p(), g, qin, qout, cin, cout,
                                                                     ctx->d_err, adj);
      else
        bounce_kernel<64, 128, SPH, SPEC, true><<<grid, 64, 0, st>>>(bvh->wnodes_p(), g, qin, qout, cin, cout,
                                                                     ctx->d_err, adj);
      return;
    }
  }
  if (ctx->tune.render_persistent) {
    // resident grid (up to 32 waves per CU); queue slots are claimed from a counter
    const int pgrid = int(std::min<int64_t>(grid, int64_t(ctx->n_cu) * 32 * 64 / THREADS));
    uint32_t* claim = ctx->d_work + 1;
    fill_async<uint32_t>(st, claim, 1, 0u);
    if (deep)
      trace_persistent<THREADS, 512, SPH><<<pgrid, THREADS, 0, st>>>(bvh->wnodes_p(), g, qin, cin, hits,
                                                                     ctx->d_err, claim);
    else
      trace_persistent<THREADS, 128, SPH><<<pgrid, THREADS, 0, st>>>(bvh->wnodes_p(), g, qin, cin, hits,
                                                                     ctx->d_err, claim);
    shade_kernel<THREADS, SPH, SPEC><<<grid, THREADS, 0, st>>>(g, qin, qout, cin, cout, hits, adj);
    return;
  }
  if (deep)
    bounce_kernel<THREADS, 512, SPH, SPEC><<<grid, THREADS, 0, st>>>(bvh->wnodes_p(), g, qin, qout, cin, cout,
                                                               ctx->d_err, adj);
  else
    bounce_kernel<THREADS, 128, SPH, SPEC><<<grid, THREADS, 0, st>>>(bvh->wnodes_p(), g, qin, qout, cin, cout,
                                                               ctx->d_err, adj);
}

template <int THREADS>
static void launch_bounce(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh, int grid,
                          const Scene* scenes, int64_t pps, int32_t k, const trace_render_params* prm,
                          int64_t path_base, float4* ro, float4* rd, float4* thr, float4* rad,
                          int32_t* qin, int32_t* qout, uint32_t* cin, uint32_t* cout, int4* hits,
                          const Adj& adj) {
  SegArgs g;
  g.tris = bvh->tris_p();
  g.perm = bvh->perm_p();
  g.scenes = scenes;
  g.pps = pps;
  g.nrm = mesh->nrm.p;
  g.mat_idx = mesh->mat_idx.p;
  g.mats = mesh->mats.p;
  g.mat_param = mesh->mat_param.p;
  g.k = k;
  g.max_depth = prm->max_depth;
  g.miss = prm->miss;
  g.path_base = path_base;
  g.ro = ro;
  g.rd = rd;
  g.thr = thr;
  g.rad = rad;
  g.seg_total = segment_counter(ctx);
  g.visits = ctx->count_visits;
  // the shader's sphere and specular branches are compiled in only for
  // meshes that need them
  const bool spec = mesh->has_specular;
  if (bvh->spheres()) {
    if (spec) launch_segments<THREADS, true, true>(ctx, bvh, grid, g, qin, qout, cin, cout, hits, adj);
    else launch_segments<THREADS, true, false>(ctx, bvh, grid, g, qin, qout, cin, cout, hits, adj);
  } else {
    if (spec) launch_segments<THREADS, false, true>(ctx, bvh, grid, g, qin, qout, cin, cout, hits, adj);
    else launch_segments<THREADS, false, false>(ctx, bvh, grid, g, qin, qout, cin, cout, hits, adj);
  }
}

namespace {

// Device state of one wavefront: per-path rays, throughput, radiance (or
// adjoint summary), two ray queues and their counters, in the context's
// scratch arena.
struct Wave {
  Scene* scenes;
  float4 *ro, *rd, *thr, *rad;
  int32_t* q[2];
  uint32_t* counts;
  int4* hits;   // hit record per queue slot (split segments only)
  int64_t spb;  // samples per batch
  // queue order (render_order): the cell-major queue, slot cells, per-tile
  // counts and their scan
  int32_t* q_sort = nullptr;
  uint8_t* q_cell = nullptr;
  uint32_t *q_counts = nullptr, *q_offsets = nullptr, *q_total = nullptr, *q_scan = nullptr;
};

int prepare(trace_ctx* ctx, const trace_bvh* const* trees, int32_t nscene, const trace_camera* cams,
            const trace_render_params* prm, Wave& w) {
  const trace_bvh* bvh = trees[0];  // every tree shares bvh's node and triangle arrays
  if (bvh->instanced) TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "render: instanced trees are traced by closest hit only");
  if (bvh->depth_bound() > 255)
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "render: tree depth %d exceeds 255", bvh->depth_bound());
  const int64_t wh = int64_t(prm->width) * prm->height;
  const int64_t per_sample = wh * nscene;
  w.spb = std::max<int64_t>(1, std::min<int64_t>(prm->spp, ctx->tune.render_batch / per_sample));
  const int64_t max_paths = w.spb * per_sample;
  // queues hold int32 path ids and uint32 counts
  if (max_paths > (int64_t(1) << 30))
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED, "render: a batch of %lld paths exceeds the 2^30 the ray queues index",
               static_cast<long long>(max_paths));
  hipStream_t st = ctx->stream;

  ScratchPlan plan;
  const size_t o_scn = plan.take(size_t(nscene) * sizeof(Scene));
  const size_t o_ro = plan.take(max_paths * sizeof(float4));
  const size_t o_rd = plan.take(max_paths * sizeof(float4));
  const size_t o_thr = plan.take(max_paths * sizeof(float4));
  const size_t o_rad = plan.take(max_paths * sizeof(float4));
  const size_t o_qa = plan.take(max_paths * sizeof(int32_t));
  const size_t o_qb = plan.take(max_paths * sizeof(int32_t));
  const size_t o_cnt = plan.take(64 * sizeof(uint32_t));
  const size_t o_hit = ctx->tune.render_persistent ? plan.take(max_paths * sizeof(int4)) : 0;
  // queue order: worth its three passes per segment once the paths are many
  // (one scene: the cells are one tree's)
  const bool qorder = !ctx->tune.render_persistent && nscene == 1 &&
                      (ctx->tune.render_order == 2 || (ctx->tune.render_order == 1 && max_paths >= kQOrderMin));
  const int64_t qblocks = ceil_div(max_paths, int64_t(kQTile));
  const size_t o_qs = qorder ? plan.take(max_paths * sizeof(int32_t)) : 0;
  const size_t o_qc = qorder ? plan.take(max_paths * sizeof(uint8_t)) : 0;
  const size_t o_qn = qorder ? plan.take(size_t(qblocks) * kQCells * sizeof(uint32_t)) : 0;
  const size_t o_qo = qorder ? plan.take(size_t(qblocks) * kQCells * sizeof(uint32_t)) : 0;
  const size_t o_qt = qorder ? plan.take(sizeof(uint32_t)) : 0;
  const size_t o_qx = qorder ? plan.take(scan::scratch_words(qblocks * kQCells) * sizeof(uint32_t)) : 0;
  int rc = ctx_reserve_scratch(ctx, plan.off);
  if (rc) return rc;
  char* base = static_cast<char*>(ctx->scratch);
  w.ro = reinterpret_cast<float4*>(base + o_ro);
  w.rd = reinterpret_cast<float4*>(base + o_rd);
  w.thr = reinterpret_cast<float4*>(base + o_thr);
  w.rad = reinterpret_cast<float4*>(base + o_rad);
  w.q[0] = reinterpret_cast<int32_t*>(base + o_qa);
  w.q[1] = reinterpret_cast<int32_t*>(base + o_qb);
  w.counts = reinterpret_cast<uint32_t*>(base + o_cnt);
  w.hits = ctx->tune.render_persistent ? reinterpret_cast<int4*>(base + o_hit) : nullptr;
  w.scenes = reinterpret_cast<Scene*>(base + o_scn);
  if (qorder) {
    w.q_sort = reinterpret_cast<int32_t*>(base + o_qs);
    w.q_cell = reinterpret_cast<uint8_t*>(base + o_qc);
    w.q_counts = reinterpret_cast<uint32_t*>(base + o_qn);
    w.q_offsets = reinterpret_cast<uint32_t*>(base + o_qo);
    w.q_total = reinterpret_cast<uint32_t*>(base + o_qt);
    w.q_scan = reinterpret_cast<uint32_t*>(base + o_qx);
  }

  // Scene table: written with kernels (arguments by value), so a render stays
  // capturable in a HIP graph with no host-memory copy node.
  for (int32_t c = 0; c < nscene; ++c) {
    Scene sc;
    const trace_camera* cam = cams + c;
    for (int a = 0; a < 3; ++a) {
      sc.cam.o[a] = float(cam->origin[a]);
      sc.cam.ll[a] = float(cam->lower_left[a]);
      sc.cam.h[a] = float(cam->horizontal[a]);
      sc.cam.v[a] = float(cam->vertical[a]);
    }
    sc.seed = prm->seed + uint64_t(c);  // scene c of a batch draws the stream of seed + c
    sc.hdr = trees[c]->header_p();
    scene_fill_kernel<<<1, 1, 0, st>>>(sc, w.scenes + c);
  }
  return TRACE_OK;
}

// ---- queue order ---------------------------------------------------------------
//
// Before each segment after the camera's, the queued paths can be regrouped by
// the cell of their ray origin (64 cells of the scene's root box, stable: a
// cell's paths keep their queue order), so that the resident waves trace paths
// that start near each other and share the nodes near them in the caches.  The
// path state stays where it is (indexed by path); neighbouring pixels' hit
// points mostly share a cell, so a cell's run of the queue keeps runs of
// consecutive paths.  On the large configuration (10.38M triangles, a tree
// larger than the Infinity Cache, 1920x1080x64 spp, depth 8) the render takes
// 102.9 ms against 107.1 ms: the scattered segments are 6-11 % faster, for
// 0.1-0.5 ms of ordering per segment.  On a 200k-triangle scene at 512x512x32
// spp (8.4M paths) the self-contained adjoint's two sweeps take 22.8 against
// 23.3 ms and the replay from kept summaries 8.5 against 8.8 ms; a 3.7M-path
// render (640x360x16 spp, 20k triangles) neither gains nor loses.  The split
// segment mode (persistent trace pass + shade pass) gains nothing from it on the
// large configuration (118.4 against 118.5 ms) and traces its queues in place.
// Cells shaped to the root box (its 6 bits given one at a time to the axis with
// the longest cells: 8 x 1 x 8 on the 2 x 0.38 x 2 terrain) rendered it in
// 102.6-102.9 ms against 102.9-103.2 with 4 x 4 x 4, within the noise: not kept.
// Each path's result is the same in any order.

__device__ inline void q_box(const Scene* __restrict__ scenes, float* box) {
  if (threadIdx.x < 3) {
    const BvhHeader* h = scenes[0].hdr;
    const float lo = h->root_lo[threadIdx.x], ext = h->root_hi[threadIdx.x] - lo;
    box[threadIdx.x] = lo;
    box[3 + threadIdx.x] = ext > 0.0f ? float(1 << kQCellBits) / ext : 0.0f;
  }
}

__device__ inline uint32_t q_cell(const float* box, float4 o) {
  const float oc[3] = {o.x, o.y, o.z};
  uint32_t key = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const uint32_t q = uint32_t(fminf(fmaxf((oc[a] - box[a]) * box[3 + a], 0.0f), float((1 << kQCellBits) - 1)));
#pragma unroll
    for (int b = 0; b < kQCellBits; ++b) key |= ((q >> b) & 1u) << (3 * b + a);
  }
  return key;
}

// Per tile of kQTile queue slots: each slot's cell (cell[i]) and the tile's
// count per cell (cell-major, tiles beyond the queue write zeros).
__global__ __launch_bounds__(kQThreads) void qcell_count_kernel(const Scene* __restrict__ scenes,
                                                               const int32_t* __restrict__ q,
                                                               const uint32_t* __restrict__ count_in,
                                                               const float4* __restrict__ ro, int32_t nblocks,
                                                               uint8_t* __restrict__ cell,
                                                               uint32_t* __restrict__ counts) {
  __shared__ uint32_t h[kQCells];
  __shared__ float box[6];
  for (int c = threadIdx.x; c < kQCells; c += kQThreads) h[c] = 0u;
  q_box(scenes, box);
  __syncthreads();
  const int64_t n = *count_in;
  const int64_t base = int64_t(blockIdx.x) * kQTile;
#pragma unroll
  for (int k = 0; k < kQPer; ++k) {
    const int64_t i = base + k * kQThreads + threadIdx.x;
    if (i < n) {
      const uint32_t c = q_cell(box, ro[q[i]]);
      cell[i] = uint8_t(c);
      atomicAdd(&h[c], 1u);
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < kQCells; c += kQThreads) counts[int64_t(c) * nblocks + blockIdx.x] = h[c];
}

// The cell-major queue: a tile's slots ranked stably by cell (one ballot per
// cell bit), staged in LDS in cell order and written as one run per cell.
__global__ __launch_bounds__(kQThreads) void qcell_scatter_kernel(const int32_t* __restrict__ q,
                                                                 const uint32_t* __restrict__ count_in,
                                                                 const uint8_t* __restrict__ cell, int32_t nblocks,
                                                                 const uint32_t* __restrict__ offsets,
                                                                 int32_t* __restrict__ q_out) {
  __shared__ int32_t sval[kQTile];
  __shared__ uint8_t scell[kQTile];
  __shared__ uint32_t slot[kQPer * kQWaves][kQCells];
  __shared__ uint32_t cstart[kQCells];
  __shared__ uint32_t coff[kQCells];
  const int tid = threadIdx.x, wave = tid / 64;
  const int64_t n = *count_in;
  const int64_t base = int64_t(blockIdx.x) * kQTile;
  if (base >= n) return;
  for (int e = tid; e < kQPer * kQWaves * kQCells; e += kQThreads) (&slot[0][0])[e] = 0u;
  __syncthreads();
  const uint64_t lt = lanemask_lt();
  uint32_t c[kQPer], r[kQPer];
#pragma unroll
  for (int k = 0; k < kQPer; ++k) {
    const int64_t i = base + k * kQThreads + tid;
    const bool valid = i < n;
    c[k] = valid ? uint32_t(cell[i]) : 0u;
    uint64_t m = ballot64(valid);
#pragma unroll
    for (int bit = 0; bit < 3 * kQCellBits; ++bit) {
      const bool set = (c[k] >> bit) & 1u;
      const uint64_t bb = ballot64(set);
      m &= set ? bb : ~bb;
    }
    r[k] = uint32_t(__popcll(m & lt));
    if (valid && (m & lt) == 0u) slot[k * kQWaves + wave][c[k]] = uint32_t(__popcll(m));
  }
  __syncthreads();
  if (tid < kQCells) {
    uint32_t run = 0;
    for (int sl = 0; sl < kQPer * kQWaves; ++sl) {
      const uint32_t v = slot[sl][tid];
      slot[sl][tid] = run;
      run += v;
    }
    cstart[tid] = run;
    coff[tid] = offsets[int64_t(tid) * nblocks + blockIdx.x];
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0;
    for (int cc = 0; cc < kQCells; ++cc) {
      const uint32_t v = cstart[cc];
      cstart[cc] = acc;
      acc += v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kQPer; ++k) {
    const int64_t i = base + k * kQThreads + tid;
    if (i < n) {
      const uint32_t p = cstart[c[k]] + slot[k * kQWaves + wave][c[k]] + r[k];
      sval[p] = q[i];
      scell[p] = uint8_t(c[k]);
    }
  }
  __syncthreads();
  const int valid_n = int(std::min<int64_t>(kQTile, n - base));
  for (int p = tid; p < valid_n; p += kQThreads) {
    const uint32_t cc = scell[p];
    q_out[int64_t(coff[cc]) + (uint32_t(p) - cstart[cc])] = sval[p];
  }
}

// Segments k = 0 .. max_depth + 1 of the paths queued in q[0] (count in counts[0]).
void bounces(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh, const Wave& w,
             int64_t npaths, int64_t pps, int64_t path_base, const trace_render_params* prm,
             const Adj& adj) {
  hipStream_t st = ctx->stream;
  const int32_t qblocks = int32_t(ceil_div(npaths, int64_t(kQTile)));
  for (int32_t k = 0; k <= prm->max_depth + 1; ++k) {
    const int in = k & 1, out = in ^ 1;
    fill_async<uint32_t>(st, w.counts + out, 1, 0u);
    int32_t* qin = w.q[in];
    if (w.q_sort && k >= 1) {
      // the scattered rays of this segment, grouped by origin cell
      qcell_count_kernel<<<qblocks, kQThreads, 0, st>>>(w.scenes, qin, w.counts + in, w.ro, qblocks, w.q_cell,
                                                        w.q_counts);
      scan::exclusive_u32(st, w.q_counts, w.q_offsets, int64_t(qblocks) * kQCells, w.q_total, w.q_scan);
      qcell_scatter_kernel<<<qblocks, kQThreads, 0, st>>>(qin, w.counts + in, w.q_cell, qblocks, w.q_offsets,
                                                          w.q_sort);
      qin = w.q_sort;
    }
    const int threads = ctx->count_visits ? 64 : ctx->tune.render_threads;
    const int grid = int(ceil_div(npaths, threads));
    if (threads == 64)
      launch_bounce<64>(ctx, mesh, bvh, grid, w.scenes, pps, k, prm, path_base, w.ro, w.rd, w.thr,
                        w.rad, qin, w.q[out], w.counts + in, w.counts + out, w.hits, adj);
    else if (threads == 256)
      launch_bounce<256>(ctx, mesh, bvh, grid, w.scenes, pps, k, prm, path_base, w.ro, w.rd, w.thr,
                         w.rad, qin, w.q[out], w.counts + in, w.counts + out, w.hits, adj);
    else
      launch_bounce<128>(ctx, mesh, bvh, grid, w.scenes, pps, k, prm, path_base, w.ro, w.rd, w.thr,
                         w.rad, qin, w.q[out], w.counts + in, w.counts + out, w.hits, adj);
  }
}

Adj mesh_adj(const trace_mesh* mesh, int32_t mode) {
  Adj a{};
  a.alb = mesh->tri_alb;
  a.emi = mesh->tri_emi;
  a.mode = mode;
  return a;
}

}  // namespace

int render(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* bvh, const trace_camera* cam,
           const trace_render_params* prm, float* d_sum, float* d_sumsq, float* d_last) {
  const trace_bvh* one[1] = {bvh};
  return render_scenes(ctx, mesh, one, 1, cam, prm, d_sum, d_sumsq, d_last, nullptr);
}

int render_scenes(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* const* trees,
                  int32_t nscene, const trace_camera* cams, const trace_render_params* prm,
                  float* d_sum, float* d_sumsq, float* d_last, float4* d_keep) {
  Wave w;
  int rc = prepare(ctx, trees, nscene, cams, prm, w);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  fill_async<unsigned long long>(st, segment_counter(ctx), 1, 0ull);
  const int64_t wh = int64_t(prm->width) * prm->height;
  // keeping summaries traces the summary sweep and derives the image from it
  const Adj adj = mesh_adj(mesh, d_keep ? kModeSummary : kModePrimal);
  for (int64_t s0 = 0; s0 < prm->spp; s0 += w.spb) {
    const int64_t nsamp = std::min<int64_t>(w.spb, prm->spp - s0);
    const int64_t pps = nsamp * wh;
    const int64_t npaths = pps * nscene;
    const int64_t path_base = (prm->sample_offset + s0) * wh;
    camera_kernel<<<int(ceil_div(npaths, 256)), 256, 0, st>>>(w.scenes, prm->width, prm->height,
                                                            npaths, pps, path_base, w.ro, w.rd,
                                                            w.thr, w.rad, w.q[0], w.counts);
    bounces(ctx, mesh, trees[0], w, npaths, pps, path_base, prm, adj);
    const bool last_batch = s0 + nsamp >= prm->spp;
    accumulate_kernel<<<int(ceil_div(wh * nscene, 256)), 256, 0, st>>>(
        w.rad, wh, nscene, int32_t(nsamp), s0 == 0 ? 1 : 0, d_sum, d_sumsq,
        last_batch ? d_last : nullptr, d_keep, int64_t(prm->spp) * wh, s0 * wh);
  }
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

int backward_scenes(trace_ctx* ctx, const trace_mesh* mesh, const trace_bvh* const* trees,
                    int32_t nscene, const trace_camera* cams, const trace_render_params* prm,
                    const float* d_grad_image, float* d_grad_albedo, float* d_grad_emission,
                    const float4* d_keep) {
  Wave w;
  int rc = prepare(ctx, trees, nscene, cams, prm, w);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  fill_async<unsigned long long>(st, segment_counter(ctx), 1, 0ull);
  fill_async<float>(st, d_grad_albedo, size_t(3 * mesh->ntri), 0.0f);
  fill_async<float>(st, d_grad_emission, size_t(3 * mesh->ntri), 0.0f);
  const int64_t wh = int64_t(prm->width) * prm->height;
  Adj summary = mesh_adj(mesh, kModeSummary);
  Adj replay = mesh_adj(mesh, kModeReplay);
  replay.g_img = d_grad_image;
  replay.g_alb = d_grad_albedo;
  replay.g_emi = d_grad_emission;
  replay.inv_spp = 1.0f / float(prm->spp);
  replay.wh = wh;
  for (int64_t s0 = 0; s0 < prm->spp; s0 += w.spb) {
    const int64_t nsamp = std::min<int64_t>(w.spb, prm->spp - s0);
    const int64_t pps = nsamp * wh;
    const int64_t npaths = pps * nscene;
    const int64_t path_base = (prm->sample_offset + s0) * wh;
    const int grid = int(ceil_div(npaths, 256));
    // sweep 1: trace every path, keep its summary (skipped when the primal
    // kept the summaries)
    if (!d_keep) {
      camera_kernel<<<grid, 256, 0, st>>>(w.scenes, prm->width, prm->height, npaths, pps, path_base,
                                          w.ro, w.rd, w.thr, w.rad, w.q[0], w.counts);
      bounces(ctx, mesh, trees[0], w, npaths, pps, path_base, prm, summary);
    }
    // sweep 2: replay the paths that carry a gradient and scatter it
    fill_async<uint32_t>(st, w.counts, 2, 0u);
    replay_camera_kernel<<<grid, 256, 0, st>>>(w.scenes, prm->width, prm->height, npaths, pps,
                                               path_base, d_grad_image, w.ro, w.rd, w.thr, w.rad,
                                               w.q[0], w.counts, d_keep, int64_t(prm->spp) * wh, s0 * wh);
    bounces(ctx, mesh, trees[0], w, npaths, pps, path_base, prm, replay);
  }
  HIP_TRY(ctx, hipGetLastError());
  return TRACE_OK;
}

}  // namespace render
