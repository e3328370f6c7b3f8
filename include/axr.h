/*
 * axr.h -- C ABI of libaxr.so, the MI355X (gfx950) rendering hot path of
 * AXI Visualizer.
 *
 * The reference drives its OpenCL kernels from Python through PyOpenCL
 * (PyOpenCL27C/Ops.py, class CLDraw); it has no native FFI.  This ABI is the
 * boundary that replaces those PyOpenCL launches: each entry point below
 * names the CLDraw method whose device work it performs (reference
 * file:line).  The Python shim axi_visualizer_amd.CLDraw keeps the
 * reference's method names and argument order and calls these through
 * ctypes.
 *
 * Conventions: plain pointers and sizes; host arrays are borrowed for the
 * call only (copied before return); readbacks fill caller-allocated arrays
 * and synchronise the context's stream.  Every call returns 0 on success or
 * a negative AXR_E* code; axr_last_error() gives the message of the last
 * failing call on the calling thread.  Calls on one context must be
 * serialised by the caller; separate contexts may be used from separate
 * threads.  Device memory lives for the context's lifetime.
 */
#ifndef AXR_H
#define AXR_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AXR_OK 0
#define AXR_EINVAL (-1)  /* bad argument */
#define AXR_EHIP (-2)    /* HIP runtime error */
#define AXR_ESTATE (-3)  /* call out of order (e.g. drawAll without clear) */
#define AXR_ECAP (-4)    /* a device capacity was exceeded */

/* Shader flags per material group (the reference selects kernels by
 * substrings of the group's shader string, Ops.py:768-805, 843-899). */
#define AXR_SH_MIP 1u    /* "mip": trilinear, av_draw_normmap_mip.c:26-123 */
#define AXR_SH_SHADOW 2u /* "shadow": PCF receive, av_alpha_max.c:53-84 */
#define AXR_SH_PHONG 4u  /* Blinn-Phong term, av_draw_normmap.c:276-310 */
#define AXR_SH_CULL 8u   /* "cull": back-face rejection, trisetup_cull.c:49-54 */

typedef struct axr_ctx axr_ctx;

const char* axr_last_error(void);

/* CLDraw.__init__(size_sky, max_uv, w, h, max_particles), Ops.py:128-229.
 * Any positive w and h.  The tiled pass's grid is w/16 x h/16 tiles
 * (Ops.py:130-131); at widths that are not multiples of 16 its work-groups
 * walk the reference's folded tile boxes (Templates.py:23-24). */
int axr_create(int w, int h, axr_ctx** out);
int axr_destroy(axr_ctx* ctx);

/* CLDraw.setScaleCull(s, cx, cy), Ops.py:231-233. */
int axr_set_scale_cull(axr_ctx* ctx, float s, float cax, float cay);

/* CLDraw.addTextureGroup(xyz, uv, vn, r, g, b, mip), Ops.py:331-378.
 * xyz, vn: [3*ntri][3] f32; uv: [3*ntri][2] f32; r, g, b: ntexels u16
 * planes; tex_size is the reference's texSize: N, or log2(N) for mips.
 * is_mip 0: N*N planes; 1: the smallest-first mip chain of
 * Utils.createMips (Utils.py:79-93); 2: the N*N base level only, the chain
 * built on the device with the same 2x2 integer box filter (ntexels N*N).
 * Returns the group index (tn) in *tn_out. */
int axr_add_texture_group(axr_ctx* ctx, const float* xyz, const float* uv, const float* vn,
                          int ntri, const uint16_t* r, const uint16_t* g, const uint16_t* b,
                          int64_t ntexels, int tex_size, int is_mip, int* tn_out);

/* Per-group material constants the reference passes as shader arguments
 * (mipBias, av_draw_normmap_mip.c:21; specPow, av_draw_normmap.c:282). */
int axr_set_group_material(axr_ctx* ctx, int tn, float mip_bias, int spec_pow);

/* CLDraw.setPrimaryLight(dirI, dirD), Ops.py:1143-1151. */
int axr_set_primary_light(axr_ctx* ctx, const float intensity[3], const float dir[3]);

/* CLDraw.vertLight(mask, dirI, dirD, pointI, pointP, spotI, spotD, spotP),
 * Ops.py:552-612 -> Pipe/vert.c:vertL.  mask[tn] != 0 selects the groups
 * lit; arrays are [n][3] f32; amb is the ambient floor (self.ambLight). */
int axr_vert_light(axr_ctx* ctx, const uint8_t* mask, int ngroups, const float* dirI,
                   const float* dirD, int ndir, const float* pointI, const float* pointP,
                   int npoint, const float* spotI, const float* spotD, const float* spotP,
                   int nspot, float amb);

/* CLDraw.setPos(vc), setVM(vM), Ops.py:475-479.  vm rows: forward,
 * horizontal, vertical (Compute.py:313-314). */
int axr_set_pos(axr_ctx* ctx, const float pos[3]);
int axr_set_vm(axr_ctx* ctx, const float vm[9]);

/* CLDraw.clearZBuffer(), Ops0.py:650-653 / Ops.py:939-946: depth to
 * 65535, colour to the clear colour (default (60,110,255)^2/3). */
int axr_clear_zbuffer(axr_ctx* ctx);
int axr_set_clear_color(axr_ctx* ctx, const uint16_t rgb[3]);

/* CLDraw.drawAll(shaders, mask, shadowIds), Ops.py:716-937 (Ops0.py:471-648).
 * shader_flags[tn]: AXR_SH_* bits; mask[tn] != 0 skips the group (the
 * reference's mask sense).  shadow_ids: shadow maps sampled by "shadow"
 * groups (at most 2).  pcf_taps: 5 (reference cross) or 9 (3x3 grid).
 * The first drawAll after axr_create needs a clearZBuffer; a later drawAll
 * with no clear before it draws onto the frame (its depth test against the
 * depth the earlier draws left), as the reference's does. */
int axr_draw_all(axr_ctx* ctx, const uint32_t* shader_flags, const uint8_t* mask, int ngroups,
                 const int* shadow_ids, int nshadow, int pcf_taps);

/* Shadow maps, Ops.py:1022-1141: addShadowMap(i, size, scale, ambLight),
 * placeShadowMap(i, pos, facing) with vm = viewMat(*facing),
 * clearShadowMap(i), shadowMap(i, whichCast, shaders, bias) (ortho). */
int axr_add_shadow_map(axr_ctx* ctx, int i, int size, float scale, float amb);
int axr_place_shadow_map(axr_ctx* ctx, int i, const float pos[3], const float vm[9], float amb);
int axr_clear_shadow_map(axr_ctx* ctx, int i);
int axr_shadow_map(axr_ctx* ctx, int i, const uint8_t* cast, int ngroups, float bias);

/* CLDraw.gamma(ex), Ops.py:947-957 -> Shaders/Post/gamma.c. */
int axr_gamma(axr_ctx* ctx, float ex);

/* CLDraw.setReflTex(name, r, g, b, size), Ops.py:458-467: a reflection
 * cube strip of faces of side `size` (size rows of 6*size texels, as
 * Cubemap.py's setupTex2 lays them out), u16 planes.  slot < 0 adds a new
 * cube map; returns its slot (>= 0) or an error code. */
int axr_set_refl_tex(axr_ctx* ctx, int slot, const uint16_t* r, const uint16_t* g,
                     const uint16_t* b, int size);
/* The cube strip in `slot` as stored on the device; returns the face size. */
int axr_get_refl_tex(axr_ctx* ctx, int slot, uint16_t* r, uint16_t* g, uint16_t* b);
/* The SSR water draw of drawAll (Ops.py:788, 859-869, 910-924 dispatch
 * groups whose shader holds "SSR" to Shaders_src/av_SSR.c):
 * groups[k] drawn with reflection cube refl[k] over the frame drawn so far
 * (the opaque drawAll).  steps = REFL_LENGTH, step = REFL_STEP, absorb =
 * ABSORB, fresnel = 0 for the metal variant's fr = 1.  The frame stays
 * HDR; getIds is not available after it. */
int axr_draw_ssr(axr_ctx* ctx, const int* groups, const int* refl, int n, int steps, int step,
                 float absorb, int fresnel);

/* CLDraw.blur(), Ops.py:1009-1020 -> Shaders/Post/blur.c (blurH) and
 * blur2.c (blurV): brightness-weighted half-resolution copy, 19-tap
 * horizontal and vertical blur, bilinear upsample added to the frame's
 * planes.  Frame must be even and at least 4x8. */
int axr_blur(axr_ctx* ctx);
/* Multi-level bloom of the frame's planes (the full-frame pipeline's post
 * pass; no single CLDraw method): `levels` fp16 half-resolution levels of
 * max(frame / white - threshold, 0), upsampled back and added with
 * weight white * strength.  W and H must be multiples of 2^levels. */
int axr_bloom(axr_ctx* ctx, int levels, float threshold, float white, float strength);

/* No counterpart in the reference.  axr_bloom's passes (same kernels, same
 * result) on caller-owned device planes r, g, b (u16, w*h each, updated in
 * place) on the caller's HIP stream (NULL: the null stream), allocating
 * nothing: for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops.  The level pyramid lives in one
 * 32 MiB scratch array per GPU, enough for frames up to 3840x2160 with 4
 * levels; larger requests fail with AXR_EINVAL.  Calls that overlap on one
 * GPU share that array and race. */
int axr_bloom_planes_noalloc(uint16_t* r, uint16_t* g, uint16_t* b, int w, int h, int levels,
                             float threshold, float white, float strength, void* stream);

/* CLDraw.getFrame(), Ops.py:1224-1230: R, G, B u16 planes and f32 depth,
 * each [h][w].  Any pointer may be NULL. */
int axr_get_frame(axr_ctx* ctx, uint16_t* r, uint16_t* g, uint16_t* b, float* depth);
/* The frame as packed RGBA8 (after gamma), [h][w] u32. */
int axr_get_rgba8(axr_ctx* ctx, uint32_t* rgba);

/* Page-locked host memory (hipHostMalloc) for readback destinations:
 * device-to-host copies into it run at the link's DMA rate, where pageable
 * destinations are staged and fault their pages in.  No counterpart in the
 * reference (its readback arrays, Ops.py:216-219, are pageable numpy
 * arrays allocated once); CLDraw keeps its getFrame/getRGBA8 arrays here. */
int axr_host_alloc(uint64_t bytes, void** out);
int axr_host_free(void* p);
/* Depth-winning global triangle index per pixel (0xFFFFFFFF: none);
 * global index = sum of earlier groups' ntri + index within the group. */
int axr_get_ids(axr_ctx* ctx, uint32_t* ids);
/* CLDraw.getSHM(i), Ops.py:1216-1222: shadow map depths, [size][size] f32. */
int axr_get_shadow_map(axr_ctx* ctx, int i, float* out);

/* Waits for the context's stream. */
int axr_sync(axr_ctx* ctx);

/* Counters: [0] small tris and [1] large tris of the last drawAll, [2]
 * shadow casters of the last shadowMap, [3] screen-bin entries of the last
 * drawAll, [4] bin entries lost since the last readback (0 unless the spill
 * list overflowed; the frame is then incomplete), [5] shadow-bin entries of
 * the last shadowMap.  Synchronises. */
int axr_get_stats(axr_ctx* ctx, uint32_t out[6]);

/* Bin bucket capacities, in entries per bin: small triangles and large-
 * triangle tiles per 32x32 screen bin, casters per 32x32 shadow bin; and
 * the spill list of each pass kind (entries past a full bucket).  0 keeps a
 * value.  Takes effect at the next pass of each kind; readbacks grow the
 * buckets after a pass spilled and the spill list after entries were lost.
 * No reference counterpart (the reference's tile bins hold 128 entries and
 * overflow silently, Pipe/coarse.c:81-82); used to tune and to test. */
int axr_set_bin_capacity(axr_ctx* ctx, uint32_t small, uint32_t large, uint32_t shadow,
                         uint32_t spill);

/* Per-kernel timing with HIP events on the context's stream.
 * axr_set_timing(ctx, 1) starts recording; axr_get_timing copies up to n
 * (name, total_ms, launches) rows and returns the row count. */
int axr_set_timing(axr_ctx* ctx, int enable);
int axr_get_timing(axr_ctx* ctx, char* names, int name_len, double* total_ms, int64_t* count,
                   int n);

/* Texels of group tn as stored on the device (r, g, b planes of the N*N
 * texture or of its whole smallest-first mip chain). */
int axr_get_texels(axr_ctx* ctx, int tn, uint16_t* r, uint16_t* g, uint16_t* b);

/* Per-vertex light of group tn (the LI buffer Pipe/vert.c:vertL writes,
 * Ops.py:605-611): out[3T][3] f32 as stored on the device. */
int axr_get_vertex_light(axr_ctx* ctx, int tn, float* out);

/* Shader binding of drawAll (flags per group, mask = groups not drawn)
 * without drawing: prepares a context used as a resident scene by batches. */
int axr_set_shaders(axr_ctx* ctx, const uint32_t* shader_flags, const uint8_t* mask,
                    int ngroups);

/* Device bytes a context holds (geometry, textures, records, bins, frame). */
int axr_get_bytes(axr_ctx* ctx, uint64_t* out);

/* ------------------------------------------------------------ batches
 * No reference counterpart.  A batch renders many jobs -- a context used as
 * a resident scene, seen by one camera at one resolution -- with one setup
 * launch, one order launch and one persistent raster launch for all of
 * them, on the batch's own stream.  Each job's pixels are bit-identical to
 * the same view drawn alone through axr_draw_all.  Batches may be in flight
 * together; a scene must not be changed while a batch that uses it runs. */
typedef struct axr_batch axr_batch;

typedef struct axr_job {
  axr_ctx* scene;         /* groups, shader flags, lights, shadow maps */
  float pos[3];           /* camera position (setPos) */
  float vm[9];            /* view basis rows (setVM) */
  float scale, cax, cay;  /* setScaleCull for this resolution */
  int width, height;      /* positive; as axr_create */
  int shadows;            /* sample the scene's first two placed shadow maps */
  int pcf_taps;           /* 5 (reference cross) or 9 (3x3) */
  float exposure;         /* gamma exposure of the RGBA8 output */
} axr_job;

int axr_batch_create(axr_batch** out);
int axr_batch_destroy(axr_batch* b);
/* Bin size 8, 16 or 32 pixels (a scheduling unit; results do not depend on
 * it), bucket capacities per bin (0: default for the size), and whether
 * the launch sequence is captured into a hipGraph and replayed while the
 * batch layout is unchanged. */
int axr_batch_configure(axr_batch* b, int bin_size, uint32_t cap_small, uint32_t cap_large,
                        int use_graph);
/* Entries of the batch's spill list (pairs of bin slot and entry taken when
 * a bucket is full), from the next render on; it grows by itself when a
 * render loses entries.  Mirrors axr_set_bin_capacity's spill argument. */
int axr_batch_set_spill_capacity(axr_batch* b, uint32_t spill);
/* Per-kernel timing of the batch (HIP events on its stream; a timed render
 * runs its three launches without the graph).  axr_batch_get_timing waits
 * for the last render and gives the setup, bin-order and raster
 * milliseconds, and the bin entries the setup appended (may be NULL). */
int axr_batch_set_timing(axr_batch* b, int enable);
int axr_batch_get_timing(axr_batch* b, float ms[3], uint32_t* bin_entries);
/* Statistics renders: while enabled, renders (without graph replay) count
 * the triangles each job's setup accepts; axr_batch_get_visible waits for
 * the last render and copies the njobs counts. */
int axr_batch_set_stats(axr_batch* b, int enable);
int axr_batch_get_visible(axr_batch* b, uint32_t* out, int njobs);
/* Enqueues the jobs; returns without waiting.  output: AXR_OUT_RGBA8
 * (gamma'd colour + depth), AXR_OUT_IDS (winner ids + depth) or
 * AXR_OUT_HDR (u16 colour planes + depth). */
#define AXR_OUT_RGBA8 0
#define AXR_OUT_IDS 1
#define AXR_OUT_HDR 2
int axr_batch_render(axr_batch* b, const axr_job* jobs, int njobs, int output);
/* 1 when the last render has finished, 0 while it runs. */
int axr_batch_query(axr_batch* b);
/* Waits for the last render; GPU time of it in *gpu_ms (may be NULL).
 * AXR_ECAP when bin entries were lost (render again: storage has grown);
 * until the next axr_batch_render every wait and every readback of that
 * render fails with AXR_ECAP, so an incomplete frame is never returned. */
int axr_batch_wait(axr_batch* b, float* gpu_ms);
int axr_batch_get_rgba8(axr_batch* b, int job, uint32_t* out);
int axr_batch_get_depth(axr_batch* b, int job, float* out);
int axr_batch_get_ids(axr_batch* b, int job, uint32_t* out);
/* HDR planes of a job of a batch rendered with AXR_OUT_HDR. */
int axr_batch_get_hdr(axr_batch* b, int job, uint16_t* r, uint16_t* g, uint16_t* bl);
/* The cube map pass of the full-frame pipeline: the scene seen from pos as
 * six size x size HDR faces (one batch of six jobs, 90 degree field of
 * view, its groups as the scene's setShaders binds them) written into
 * reflection cube slot `slot` (< 0: a new slot) in the strip layout of
 * setReflTex that the SSR shader samples (Shaders_src/av_SSR.c:107-128):
 * face f spans columns [f*size, (f+1)*size).  The engine's makecubemap
 * (3SSR.py:67-76) asks for this through a drawCubeMap it never defines.
 * Returns the slot.  The scene's later draws are ordered after it. */
int axr_render_cube(axr_batch* b, axr_ctx* scene, int slot, const float pos[3], int size,
                    int shadows, int pcf_taps);
/* Device bytes the batch holds (records, bins, outputs). */
int axr_batch_get_bytes(axr_batch* b, uint64_t* out);

/* Self-check of the device arithmetic the raster kernels use beyond plain
 * IEEE operations (no counterpart in the reference, whose Python/numpy
 * arithmetic divides in IEEE fp32 and converts with saturating casts): on
 * the current device, compares the Newton reciprocal with IEEE 1/b for
 * every b with |b| in [2^-60, 2^60], the Markstein quotient with IEEE a/b
 * on rounds x 2^26 random pairs from the ranges the kernels admit, and the
 * hardware's saturating float-to-int conversions with their formulas for
 * every float.  out[0] = reciprocal mismatches, out[1] = pairs tested,
 * out[2] = quotient mismatches, out[3] = conversion mismatches; the
 * mismatch counts must all be 0. */
int axr_check_arith(uint64_t seed, uint32_t rounds, uint64_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
