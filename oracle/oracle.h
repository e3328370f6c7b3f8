/*
 * oracle.h -- CPU restatement of the AXI Visualizer rendering hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP path in axi_visualizer_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it; the product never does.
 *
 * It restates, in serial order and with FP contraction off, the OpenCL
 * kernels of the reference (paths relative to the reference's PyOpenCL27C/):
 *   Pipe/trisetup.c, Pipe/trisetup_cull.c:49-54   -> orc_trisetup
 *   Pipe/coarse.c                                 -> orc_coarse_tri
 *   Shaders_src/Templates.py:200-326 (drawSmall)  -> raster_small (internal)
 *   Shaders_src/Templates.py:4-191   (tiled draw) -> raster_tile  (internal)
 *   Shaders_src/av_draw_base.c, av_draw_normmap_mip.c:26-123,
 *   av_draw_normmap.c:276-310                     -> shade (internal)
 *   Shaders/drawreflcubemapfresneltexsh.c:162-182,
 *   Shaders_src/av_alpha_max.c:53-84 (PCF)        -> pcf (internal)
 *   Pipe/trisetupOrtho.c, Shaders/drawmin.c       -> shadow pass
 *   Shaders/clearzb.c, Shaders/Post/gamma.c       -> clear, orc_gamma
 *   Shaders/Post/blur.c, Post/blur2.c, Ops.py:1009 -> orc_blur
 *   Pipe/vert.c:38-88                              -> orc_vertlight
 *   Utils.py:79-93 (createMips)                    -> orc_create_mips
 *   Ops0.py:471-648, Ops.py:1089-1141 (pass order) -> orc_render
 *
 * Arithmetic decrees where OpenCL leaves freedom (same on every path):
 *   IEEE-754 binary32, round-to-nearest-even, no contraction, IEEE division
 *   and sqrt; dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z; min(x,y) = y<x?y:x;
 *   max(x,y) = x<y?y:x; float->int conversion truncates and saturates
 *   (NaN -> 0); float->ushort saturates; normalize(v) = v * (1/sqrt(v.v));
 *   pown by binary exponentiation.
 *
 * Depth-winner semantics (the reference races; this fixes it): passes are
 * ordered (all groups' small triangles, then each group's large triangles),
 * triangles ascend within a pass, a fragment writes iff depth >= tz, so the
 * final depth is the minimum and the winner is the last fragment at that
 * minimum.
 */
#ifndef AXR_ORACLE_H
#define AXR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MIP 1u    /* trilinear from a smallest-first mip chain */
#define ORC_SHADOW 2u /* receive PCF shadows */
#define ORC_PHONG 4u  /* Blinn-Phong specular term */
#define ORC_CULL 8u   /* trisetup_cull back-face rejection */

typedef struct {
  int ntri;
  const float* xyz; /* [3*ntri][3] world positions */
  const float* vn;  /* [3*ntri][3] vertex normals (reference sign) */
  const float* uv;  /* [3*ntri][2] */
  const float* li;  /* [3*ntri][3] per-vertex light I */
  const uint16_t* tr; /* texture planes: N*N, or the mip chain */
  const uint16_t* tg;
  const uint16_t* tb;
  int tex_size;       /* N (bilinear) or log2(N) (ORC_MIP), as texSize */
  uint32_t flags;
  float mip_bias;
  int spec_pow;
  int skip; /* drawAll mask: 1 = group not drawn */
  int cast; /* casts into shadow maps */
  int ssr;  /* >= 0: an SSR water group (av_SSR.c) reflecting cube map
               refl_*[ssr]; drawn after the opaque groups; -1 otherwise */
} OrcGroup;

typedef struct {
  float pos[3];
  float vm[9]; /* rows: forward, horizontal, vertical */
  float scale;
  float cax, cay;
} OrcCam;

typedef struct {
  int dim;
  OrcCam cam; /* pos, vm, scale used; cax/cay ignored */
  float bias;
} OrcShadowCfg;

typedef struct {
  int W, H;
  OrcCam view;
  float lint[3];
  float ldir[3];
  float amb;
  int nshadow; /* 0..2 cascades */
  OrcShadowCfg shadow[2];
  int pcf_taps;   /* 5 = reference cross, 9 = 3x3 */
  float exposure; /* > 0: gamma to 8 bit after drawing */
  uint16_t clear_rgb[3];
  int nthreads; /* 0 = all cores */
  /* reflection cube strips (setReflTex): faces of side refl_size[k], as
   * refl_size rows of 6*refl_size texels per plane */
  const uint16_t* refl_r[4];
  const uint16_t* refl_g[4];
  const uint16_t* refl_b[4];
  int refl_size[4];
  int ssr_steps, ssr_step; /* REFL_LENGTH, REFL_STEP */
  float ssr_absorb;        /* ABSORB */
  int ssr_fresnel;         /* 0: fr = 1 (metal variant) */
  /* optional: the frame this drawAll draws onto (no clearZBuffer before it,
   * Ops.py:716-946): depth, winner ids and colour planes as a previous
   * render left them (exposure 0); null: start from the clear */
  const float* init_depth;
  const uint32_t* init_id;
  const uint16_t* init_r;
  const uint16_t* init_g;
  const uint16_t* init_b;
} OrcFrame;

typedef struct {
  float* depth;  /* [H*W] */
  uint32_t* id;  /* [H*W] winner global triangle index, 0xFFFFFFFF none */
  uint16_t* r;   /* [H*W] colour planes */
  uint16_t* g;
  uint16_t* b;
  float* shadow[2]; /* optional [dim*dim] copies of the shadow maps */
  int64_t stats[8]; /* 0 small, 1 large, 2 fragments, 3 max bin, 4 bin entries,
                       5/6 shadow fragments of cascade 0/1 */
} OrcOut;

int orc_render(const OrcGroup* groups, int ngroups, const OrcFrame* fr,
               OrcOut* out);

/* Per-triangle setup. cls: 0 rejected, 1 small, 2 large. */
int orc_trisetup(const float* xyz, const float* vn, int ntri, const OrcCam* cam,
                 int W, int H, int cull, float* P, float* Z, int8_t* cls);

/* Ortho shadow setup: XY int pairs, Z, ok flag per triangle. */
int orc_shadow_setup(const float* xyz, int ntri, const OrcCam* cam, int dim,
                     float bias, int32_t* XY, float* Z, int8_t* ok);

/* Coarse tile list (duplicates kept, in walk order). Returns the count,
 * or -(count) when cap was too small. */
int orc_coarse_tri(const float* P6, int WC, int HC, int32_t* tiles, int cap);

/* Fragments of one triangle in evaluation order.  mode 0: small path over a
 * W x H screen; mode 1: tiled path restricted to tile index `tile`;
 * mode 2: shadow depth path (P6 holds int coordinates as floats).
 * Writes x, y, tz, half; returns the count or -(count) if cap too small. */
int orc_fragments(int mode, const float* P6, const float* Z3, int W, int H,
                  int tile, int32_t* xs, int32_t* ys, float* tz, int8_t* half,
                  int cap);

int orc_gamma(uint16_t* r, uint16_t* g, uint16_t* b, int64_t n, float ex);
/* CLDraw.blur: Post/blur.c blurH then Post/blur2.c blurV on u16 planes of
 * a W x H frame (serial phases; decrees in oracle.cpp). */
int orc_blur(uint16_t* r, uint16_t* g, uint16_t* b, int W, int H);
/* The oracle's decreed elementary functions (which 0: log2, 1: exp), as the
 * kernels compute them, over n floats. */
int orc_math(int which, const float* x, float* y, int64_t n);
int orc_create_mips(const uint16_t* plane, int N, uint16_t* out);
int orc_vertlight(const float* xyz, const float* vn, int nvert,
                  const float* dirI, const float* dirD, int ndir,
                  const float* pointI, const float* pointP, int npoint,
                  const float* spotI, const float* spotD, const float* spotP,
                  int nspot, float amb, float* out);

#ifdef __cplusplus
}
#endif
#endif
