// batch_api.cpp -- C ABI of batched rendering (declared in include/axr.h).
//
// A batch renders jobs -- a resident scene (an axr_ctx holding geometry,
// textures, lights and shadow maps) seen by one camera at one resolution --
// with three launches on the batch's own stream (batch.hip): setup of every
// (job, triangle), one order pass over every job's bins, and one persistent
// raster that drains the heavy-first bin queue.  Jobs may mix scenes and
// resolutions.  The launch sequence can be captured once into a hipGraph
// and replayed while the layout (job count, sizes, scenes, bin size) stays
// the same; cameras are re-uploaded before every replay.
//
// Several batches may be in flight at once (each has its own stream and
// buffers); that is how the render server overlaps batches on the device.
#include "axr_ctx.h"

namespace {

constexpr uint32_t kBatchSpill = 1u << 20;
// (job, triangle) pairs of large triangles a render can list before its
// coarse pass; grown like the spill list when a render loses some
constexpr uint32_t kBatchLarge = 1u << 16;

struct JobInfo {
  axr_ctx* scene;
  int W, H;
  size_t out_off;  // pixel offset of this job in the output pools
};

// Default bucket capacity per bin for a bin of 2^shift pixels on a side:
// the single-frame capacities scaled by bin area.
uint32_t default_cap(int shift, bool large) {
  const uint32_t base = large ? 1024u : 4096u;  // for 32-px bins
  const int scale = 2 * (5 - shift);
  uint32_t c = base >> scale;
  return c < 64u ? 64u : c;
}

}  // namespace

struct axr_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  int bin_shift = 5;
  int occupancy[3][3] = {};  // raster blocks per CU by (bin shift - 3, shade mode)
  BinStore bins;
  uint32_t* d_head = nullptr;
  uint2* d_large = nullptr;        // large triangles of the render: (job, triangle)
  uint32_t* d_large_ctl = nullptr;  // [0]: pairs listed (reset by the order pass)
  uint32_t large_cap = 0, want_large = kBatchLarge;
  OrderScratch* d_order = nullptr;  // the many-block order pass's histogram and cursors
  uint32_t* d_counters = nullptr;
  // pinned copy of d_counters, written on the stream after every render so
  // that a wait never issues a copy of its own: a synchronous copy would
  // queue behind whatever other batch holds the device
  uint32_t* h_ctr = nullptr;
  std::vector<hipEvent_t> scene_ev;  // reused: one per distinct scene of a render
  AxrTriRec* d_rec = nullptr;
  size_t rec_cap = 0;
  float* d_depth = nullptr;
  uint32_t* d_rgba = nullptr;
  uint32_t* d_ids = nullptr;
  size_t px_cap = 0;
  uint16_t* d_hdr = nullptr;  // HDR planes [3][hdr_cap], for HDR renders
  size_t hdr_cap = 0;
  AxrJobDev* d_jobs = nullptr;
  AxrJobDev* h_jobs = nullptr;  // pinned staging of the job table
  // setup groups, one per scene of the batch: uint4 (first setup block,
  // offset of its jobs in the index list, job count, triangles) x njobs at
  // most, then the job index list (njobs words); pinned staging beside it
  uint32_t* d_sg = nullptr;
  uint32_t* h_sg = nullptr;
  int nsg = 0;
  int jobs_cap = 0;
  std::vector<JobInfo> info;
  int njobs = 0;
  uint32_t setup_blocks = 0;
  int nb_total = 0;
  int mode = SHADE_LDR8;  // ShadeMode of the last render
  hipEvent_t ev_start = nullptr, ev_done = nullptr;
  hipEvent_t ev_copied = nullptr;  // the job table has left the staging buffer
  bool rendered = false;
  // the last render lost bin entries: its frames are incomplete until the
  // next render (every wait and readback of it fails with AXR_ECAP)
  bool incomplete = false;
  // per-kernel timing (axr_batch_set_timing): events around setup, order
  // and raster of the last render, which then runs without its graph
  bool timing = false;
  hipEvent_t ev_k[4] = {nullptr, nullptr, nullptr, nullptr};
  bool timed = false;  // the last render recorded ev_k
  // statistics renders (axr_batch_set_stats): visible triangles per job
  bool stats = false;
  uint32_t* d_visible = nullptr;
  int visible_cap = 0;
  // graph replay
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint64_t> graph_key;
};

namespace {

void free_graph(axr_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->graph_key.clear();
}

int grow_pools(axr_batch* b, size_t tris, size_t px, int njobs, bool hdr) {
  int rc;
  if (tris > b->rec_cap) {
    HIPCHK(hipStreamSynchronize(b->stream));
    dfree(b->d_rec);
    if ((rc = dalloc(&b->d_rec, tris))) return rc;
    b->rec_cap = tris;
    free_graph(b);
  }
  if (px > b->px_cap) {
    HIPCHK(hipStreamSynchronize(b->stream));
    dfree(b->d_depth);
    dfree(b->d_rgba);
    dfree(b->d_ids);
    if ((rc = dalloc(&b->d_depth, px)) || (rc = dalloc(&b->d_rgba, px)) ||
        (rc = dalloc(&b->d_ids, px)))
      return rc;
    b->px_cap = px;
    free_graph(b);
  }
  if (hdr && px > b->hdr_cap) {
    HIPCHK(hipStreamSynchronize(b->stream));
    dfree(b->d_hdr);
    if ((rc = dalloc(&b->d_hdr, 3 * px))) return rc;
    b->hdr_cap = px;
    free_graph(b);
  }
  if (njobs > b->jobs_cap) {
    HIPCHK(hipStreamSynchronize(b->stream));
    dfree(b->d_jobs);
    dfree(b->d_sg);
    if (b->h_jobs) (void)hipHostFree(b->h_jobs);
    if (b->h_sg) (void)hipHostFree(b->h_sg);
    b->h_jobs = nullptr;
    b->h_sg = nullptr;
    if ((rc = dalloc(&b->d_jobs, (size_t)njobs)) || (rc = dalloc(&b->d_sg, 5 * (size_t)njobs)))
      return rc;
    HIPCHK(hipHostMalloc((void**)&b->h_jobs, sizeof(AxrJobDev) * (size_t)njobs));
    HIPCHK(hipHostMalloc((void**)&b->h_sg, sizeof(uint32_t) * 5 * (size_t)njobs));
    b->jobs_cap = njobs;
    free_graph(b);
  }
  return 0;
}

// Device job descriptor of job j (scene state as it is now).
AxrJobDev make_job(const axr_batch* b, const axr_job& j, const JobInfo& in, uint32_t slot0,
                   uint32_t blk0, int nbx, size_t rec_off) {
  axr_ctx* sc = j.scene;
  AxrJobDev d{};
  AxrDrawConst& dc = d.dc;
  for (int k = 0; k < 3; k++) dc.view.pos[k] = j.pos[k];
  for (int k = 0; k < 9; k++) dc.view.vm[k] = j.vm[k];
  dc.view.scale = j.scale;
  dc.view.cax = j.cax;
  dc.view.cay = j.cay;
  dc.cull = make_cull(dc.view);
  for (int k = 0; k < 3; k++) {
    dc.lint[k] = sc->lint[k];
    dc.ldir[k] = sc->ldir[k];
    dc.clear_rgb[k] = sc->clear_rgb[k];
  }
  dc.amb = sc->amb;
  dc.W = in.W;
  dc.H = in.H;
  dc.WC = in.W / 16;
  dc.HC = in.H / 16;
  dc.T = sc->T;
  dc.ngroups = (uint32_t)sc->groups.size();
  dc.pcf_taps = j.pcf_taps;
  dc.flag_mask = j.shadows ? 0xFFFFFFFFu : ~AXR_SH_SHADOW;
  dc.nshadow = 0;
  if (j.shadows) {
    for (int i = 0; i < AXR_MAX_SHADOWS && dc.nshadow < AXR_SHADOW_BIND; i++) {
      const HostShadow& s = sc->sm[i];
      if (!s.exists || !s.placed) continue;
      AxrShadowDev& sd = dc.sh[dc.nshadow++];
      sd.map = s.d_map;
      sd.dim = s.dim;
      sd.wS = s.dim / 2;
      sd.cam = s.cam;
    }
  }
  d.vpos = sc->d_vpos;
  d.vattr = sc->d_vattr;
  d.groups = sc->d_groups;
  d.chunks = sc->d_chunks;
  d.rec = b->d_rec + rec_off;
  d.depth = b->d_depth + in.out_off;
  d.rgba = b->d_rgba + in.out_off;
  d.ids = b->d_ids + in.out_off;
  if (b->d_hdr && b->mode == SHADE_HDR) {
    d.r = b->d_hdr + in.out_off;
    d.g = b->d_hdr + b->hdr_cap + in.out_off;
    d.b = b->d_hdr + 2 * b->hdr_cap + in.out_off;
  }
  d.exposure = j.exposure;
  d.slot0 = slot0;
  d.blk0 = blk0;
  d.nbx = nbx;
  return d;
}

int enqueue(axr_batch* b) {
  AxrBins B = b->bins.B;
  B.head = b->d_head;
  int& occ = b->occupancy[b->bin_shift - 3][b->mode];
  if (!occ) occ = batch_raster_blocks_per_cu(b->bin_shift, b->mode);
  const int grid_max = b->cu_count * occ;
  const int grid = b->nb_total < grid_max ? b->nb_total : grid_max;
  if (b->timing) HIPCHK(hipEventRecord(b->ev_k[0], b->stream));
  HIPCHK(launch_batch_setup(b->bin_shift, b->d_jobs, (const uint4*)b->d_sg, b->nsg,
                            b->d_sg + 4 * (size_t)b->njobs, b->setup_blocks, B, b->d_large,
                            b->d_large_ctl, b->large_cap, b->d_counters,
                            b->stats ? b->d_visible : nullptr, b->stream));
  HIPCHK(launch_batch_coarse(b->bin_shift, b->d_jobs, b->d_large, b->d_large_ctl, b->large_cap,
                             B, b->cu_count * 2, b->stream));
  if (b->timing) HIPCHK(hipEventRecord(b->ev_k[1], b->stream));
  HIPCHK(launch_bin_order_wide(B, b->d_order, b->d_counters, b->d_large_ctl, b->stream));
  if (b->timing) HIPCHK(hipEventRecord(b->ev_k[2], b->stream));
  HIPCHK(launch_batch_raster(b->bin_shift, b->mode, b->d_jobs, b->njobs, B, grid, b->stream));
  if (b->timing) HIPCHK(hipEventRecord(b->ev_k[3], b->stream));
  b->timed = b->timing;
  return 0;
}

}  // namespace

extern "C" {

int axr_batch_create(axr_batch** out) {
  if (!out) return fail(AXR_EINVAL, "null out pointer");
  *out = nullptr;
  axr_batch* b = new axr_batch();
  auto bail = [&](int r) {
    axr_batch_destroy(b);
    return r;
  };
  if (hipGetDevice(&b->device) != hipSuccess) return bail(fail(AXR_EHIP, "no device"));
  if (hipDeviceGetAttribute(&b->cu_count, hipDeviceAttributeMultiprocessorCount, b->device) !=
      hipSuccess)
    return bail(fail(AXR_EHIP, "device attribute query failed"));
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev_start) != hipSuccess || hipEventCreate(&b->ev_done) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_copied, hipEventDisableTiming) != hipSuccess)
    return bail(fail(AXR_EHIP, "stream/event creation failed"));
  int rc;
  if ((rc = dalloc(&b->d_head, 1)) || (rc = dalloc(&b->d_counters, AXR_CTR_WORDS)) ||
      (rc = dalloc(&b->d_large_ctl, 2)) || (rc = dalloc(&b->d_order, 1)))
    return bail(rc);
  if (hipHostMalloc((void**)&b->h_ctr, sizeof(uint32_t) * AXR_CTR_WORDS) != hipSuccess)
    return bail(fail(AXR_EHIP, "pinned counter allocation failed"));
  std::memset(b->h_ctr, 0, sizeof(uint32_t) * AXR_CTR_WORDS);
  if (launch_fill_u32(b->d_counters, AXR_CTR_WORDS, 0, b->stream) != hipSuccess ||
      launch_fill_u32(b->d_head, 1, 0, b->stream) != hipSuccess ||
      launch_fill_u32(b->d_large_ctl, 2, 0, b->stream) != hipSuccess ||
      launch_fill_u32((uint32_t*)b->d_order, sizeof(OrderScratch) / sizeof(uint32_t), 0,
                      b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return bail(fail(AXR_EHIP, "initial fill failed"));
  b->bins.nseg = 2;
  b->bins.want_spill = kBatchSpill;
  b->bins.want0 = default_cap(b->bin_shift, false);
  b->bins.want1 = default_cap(b->bin_shift, true);
  *out = b;
  return AXR_OK;
}

int axr_batch_destroy(axr_batch* b) {
  if (!b) return AXR_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  free_graph(b);
  bins_free(b->bins);
  dfree(b->d_head);
  dfree(b->d_large);
  dfree(b->d_large_ctl);
  dfree(b->d_order);
  dfree(b->d_visible);
  dfree(b->d_counters);
  dfree(b->d_rec);
  dfree(b->d_depth);
  dfree(b->d_rgba);
  dfree(b->d_ids);
  dfree(b->d_hdr);
  dfree(b->d_jobs);
  dfree(b->d_sg);
  if (b->h_jobs) (void)hipHostFree(b->h_jobs);
  if (b->h_sg) (void)hipHostFree(b->h_sg);
  if (b->h_ctr) (void)hipHostFree(b->h_ctr);
  for (hipEvent_t e : b->scene_ev) (void)hipEventDestroy(e);
  if (b->ev_start) (void)hipEventDestroy(b->ev_start);
  if (b->ev_done) (void)hipEventDestroy(b->ev_done);
  if (b->ev_copied) (void)hipEventDestroy(b->ev_copied);
  for (hipEvent_t e : b->ev_k)
    if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return AXR_OK;
}

int axr_batch_configure(axr_batch* b, int bin_size, uint32_t cap_small, uint32_t cap_large,
                        int use_graph) {
  if (!b) return fail(AXR_EINVAL, "null batch");
  int shift;
  if (bin_size == 8)
    shift = 3;
  else if (bin_size == 16)
    shift = 4;
  else if (bin_size == 32)
    shift = 5;
  else
    return fail(AXR_EINVAL, "bin size must be 8, 16 or 32 (got %d)", bin_size);
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->bin_shift = shift;
  b->bins.want0 = cap_small ? cap_small : default_cap(shift, false);
  b->bins.want1 = cap_large ? cap_large : default_cap(shift, true);
  b->use_graph = use_graph != 0;
  free_graph(b);
  return AXR_OK;
}

int axr_batch_set_timing(axr_batch* b, int enable) {
  if (!b) return fail(AXR_EINVAL, "null batch");
  HIPCHK(hipSetDevice(b->device));
  for (hipEvent_t& e : b->ev_k)
    if (!e) HIPCHK(hipEventCreate(&e));
  b->timing = enable != 0;
  return AXR_OK;
}

int axr_batch_get_timing(axr_batch* b, float ms[3], uint32_t* bin_entries) {
  if (!b || !ms) return fail(AXR_EINVAL, "batch timing: bad arguments");
  int rc;
  if ((rc = axr_batch_wait(b, nullptr))) return rc;
  if (!b->timed) return fail(AXR_ESTATE, "the last render was not timed (axr_batch_set_timing)");
  for (int k = 0; k < 3; k++) HIPCHK(hipEventElapsedTime(&ms[k], b->ev_k[k], b->ev_k[k + 1]));
  if (bin_entries) *bin_entries = b->h_ctr[AXR_CTR_BIN_TOTAL];
  return AXR_OK;
}

int axr_batch_set_stats(axr_batch* b, int enable) {
  if (!b) return fail(AXR_EINVAL, "null batch");
  b->stats = enable != 0;
  return AXR_OK;
}

int axr_batch_get_visible(axr_batch* b, uint32_t* out, int njobs) {
  if (!b || !out) return fail(AXR_EINVAL, "batch visible: bad arguments");
  if (!b->stats || !b->d_visible || njobs != b->njobs)
    return fail(AXR_ESTATE, "batch visible: the last render was not a statistics render of %d jobs",
                njobs);
  int rc;
  if ((rc = axr_batch_wait(b, nullptr))) return rc;
  HIPCHK(hipMemcpy(out, b->d_visible, sizeof(uint32_t) * (size_t)njobs, hipMemcpyDeviceToHost));
  return AXR_OK;
}

int axr_batch_set_spill_capacity(axr_batch* b, uint32_t spill) {
  if (!b) return fail(AXR_EINVAL, "null batch");
  if (spill == 0) return fail(AXR_EINVAL, "spill capacity must be positive");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->bins.want_spill = spill;
  free_graph(b);
  return AXR_OK;
}

int axr_batch_render(axr_batch* b, const axr_job* jobs, int njobs, int output) {
  if (!b || (!jobs && njobs)) return fail(AXR_EINVAL, "batch render: bad arguments");
  if (njobs < 0) return fail(AXR_EINVAL, "negative job count");
  if (output != AXR_OUT_RGBA8 && output != AXR_OUT_IDS && output != AXR_OUT_HDR)
    return fail(AXR_EINVAL, "batch render: output %d", output);
  HIPCHK(hipSetDevice(b->device));
  for (int i = 0; i < njobs; i++) {
    const axr_job& j = jobs[i];
    if (!j.scene) return fail(AXR_EINVAL, "job %d: null scene", i);
    if (j.width <= 0 || j.height <= 0)
      return fail(AXR_EINVAL, "job %d: width and height must be positive", i);
    if ((int64_t)j.width * j.height > (int64_t)1 << 28)
      return fail(AXR_EINVAL, "job %d: frame too large", i);
    if (j.pcf_taps != 5 && j.pcf_taps != 9) return fail(AXR_EINVAL, "job %d: pcf_taps 5 or 9", i);
    if (j.scene->device != b->device) return fail(AXR_EINVAL, "job %d: scene on another device", i);
  }
  int rc;
  // Scenes: run pending shadow passes and clears, then order the batch
  // stream after the scene's stream.
  std::vector<axr_ctx*> scenes;
  for (int i = 0; i < njobs; i++) {
    axr_ctx* sc = jobs[i].scene;
    bool seen = false;
    for (auto* s : scenes) seen |= s == sc;
    if (seen) continue;
    scenes.push_back(sc);
    if ((rc = flush_shadows(sc))) return rc;
    for (int m = 0; m < AXR_MAX_SHADOWS; m++)
      if (sc->sm[m].exists && sc->sm[m].placed && (rc = apply_shadow_clear(sc, m))) return rc;
    if (scenes.size() > b->scene_ev.size()) {
      hipEvent_t ev;
      HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      b->scene_ev.push_back(ev);
    }
    hipEvent_t ev = b->scene_ev[scenes.size() - 1];
    HIPCHK(hipEventRecord(ev, sc->stream));
    HIPCHK(hipStreamWaitEvent(b->stream, ev, 0));
  }
  // Layout: bins, record and pixel offsets in job order; setup blocks per
  // scene (the jobs of one scene share its setup blocks, in job order).
  const int bin = 1 << b->bin_shift;
  b->info.assign((size_t)njobs, JobInfo{});
  size_t tris = 0, px = 0;
  uint32_t slots = 0, blocks = 0;
  std::vector<uint32_t> slot0((size_t)njobs), blk0((size_t)njobs);
  std::vector<int> nbx((size_t)njobs);
  std::vector<size_t> rec_off((size_t)njobs);
  for (int i = 0; i < njobs; i++) {
    const axr_job& j = jobs[i];
    JobInfo& in = b->info[(size_t)i];
    in.scene = j.scene;
    in.W = j.width;
    in.H = j.height;
    in.out_off = px;
    px += (size_t)j.width * j.height;
    nbx[(size_t)i] = (j.width + bin - 1) / bin;
    slot0[(size_t)i] = slots;
    slots += (uint32_t)(nbx[(size_t)i] * ((j.height + bin - 1) / bin));
    rec_off[(size_t)i] = tris;
    tris += j.scene->T;
  }
  // scene groups, in order of first appearance (scenes lists them so)
  std::vector<uint32_t> grp_jobs;  // job indices grouped by scene
  std::vector<uint32_t> grp_desc;  // per group: blk0, first index, count, T
  grp_jobs.reserve((size_t)njobs);
  for (axr_ctx* sc : scenes) {
    const uint32_t first = (uint32_t)grp_jobs.size();
    for (int i = 0; i < njobs; i++) {
      if (jobs[i].scene != sc) continue;
      grp_jobs.push_back((uint32_t)i);
      blk0[(size_t)i] = blocks;
    }
    const uint32_t nblk = (sc->T + 255u) / 256u;
    if (nblk == 0) continue;  // no triangles: nothing to set up
    grp_desc.insert(grp_desc.end(), {blocks, first, (uint32_t)grp_jobs.size() - first, sc->T});
    blocks += nblk;
  }
  b->njobs = njobs;
  b->nsg = (int)(grp_desc.size() / 4);
  b->setup_blocks = blocks;
  b->nb_total = (int)slots;
  b->mode = output == AXR_OUT_IDS ? SHADE_IDS : output == AXR_OUT_HDR ? SHADE_HDR : SHADE_LDR8;
  if ((rc = grow_pools(b, tris ? tris : 1, px ? px : 1, njobs ? njobs : 1,
                       b->mode == SHADE_HDR)))
    return rc;
  if (b->stats && njobs > b->visible_cap) {
    HIPCHK(hipStreamSynchronize(b->stream));
    dfree(b->d_visible);
    b->visible_cap = 0;
    if ((rc = dalloc(&b->d_visible, (size_t)njobs))) return rc;
    b->visible_cap = njobs;
  }
  if (b->stats) HIPCHK(launch_fill_u32(b->d_visible, (size_t)njobs, 0, b->stream));
  if (b->large_cap < b->want_large) {
    HIPCHK(hipStreamSynchronize(b->stream));
    dfree(b->d_large);
    b->large_cap = 0;
    if ((rc = dalloc(&b->d_large, b->want_large))) return rc;
    b->large_cap = b->want_large;
    free_graph(b);
  }
  const AxrBins before = b->bins.B;
  if ((rc = bins_prepare(b->bins, slots ? (int)slots : 1, b->stream))) return rc;
  b->bins.B.nb = (int)slots;
  if (before.ent != b->bins.B.ent || before.cnt != b->bins.B.cnt ||
      before.spill != b->bins.B.spill)
    free_graph(b);
  // the previous render's table copy must be done before the staging is rewritten
  if (b->rendered) HIPCHK(hipEventSynchronize(b->ev_copied));
  for (int i = 0; i < njobs; i++)
    b->h_jobs[i] = make_job(b, jobs[i], b->info[(size_t)i], slot0[(size_t)i], blk0[(size_t)i],
                            nbx[(size_t)i], rec_off[(size_t)i]);
  std::memcpy(b->h_sg, grp_desc.data(), grp_desc.size() * sizeof(uint32_t));
  std::memcpy(b->h_sg + 4 * (size_t)njobs, grp_jobs.data(), grp_jobs.size() * sizeof(uint32_t));
  b->rendered = true;
  b->incomplete = false;
  HIPCHK(hipEventRecord(b->ev_start, b->stream));
  if (njobs) {
    HIPCHK(hipMemcpyAsync(b->d_jobs, b->h_jobs, sizeof(AxrJobDev) * (size_t)njobs,
                          hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d_sg, b->h_sg, sizeof(uint32_t) * 5 * (size_t)njobs,
                          hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipEventRecord(b->ev_copied, b->stream));
    b->timed = false;
    if (b->use_graph && !b->timing && !b->stats) {
      // The graph holds the three launches; its arguments depend on this key.
      std::vector<uint64_t> key = {(uint64_t)njobs, (uint64_t)b->nsg, blocks, slots,
                                   (uint64_t)b->bin_shift,
                                   (uint64_t)b->mode, b->bins.B.cap0, b->bins.B.cap1,
                                   (uint64_t)(uintptr_t)b->bins.B.ent,
                                   (uint64_t)(uintptr_t)b->d_jobs,
                                   (uint64_t)(uintptr_t)b->d_sg};
      if (!b->gexec || key != b->graph_key) {
        free_graph(b);
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue(b);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        if (e != hipSuccess) return fail(AXR_EHIP, "graph capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(AXR_EHIP, "graph instantiate: %s", hipGetErrorString(e));
        b->graph_key = key;
      }
      HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    } else if ((rc = enqueue(b))) {
      return rc;
    }
    HIPCHK(hipMemcpyAsync(b->h_ctr, b->d_counters, sizeof(uint32_t) * AXR_CTR_WORDS,
                          hipMemcpyDeviceToHost, b->stream));
  }
  HIPCHK(hipEventRecord(b->ev_done, b->stream));
  return AXR_OK;
}

int axr_batch_query(axr_batch* b) {
  if (!b) return fail(AXR_EINVAL, "null batch");
  const hipError_t e = hipEventQuery(b->ev_done);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return fail(AXR_EHIP, "batch query: %s", hipGetErrorString(e));
}

int axr_batch_wait(axr_batch* b, float* gpu_ms) {
  if (!b) return fail(AXR_EINVAL, "null batch");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipEventSynchronize(b->ev_done));
  if (gpu_ms) {
    *gpu_ms = 0.0f;
    if (b->rendered) HIPCHK(hipEventElapsedTime(gpu_ms, b->ev_start, b->ev_done));
  }
  // the counters as the last render left them (copied on the stream before
  // ev_done); spill and loss accumulate on the device until seen here
  uint32_t h[AXR_CTR_WORDS];
  std::memcpy(h, (const void*)b->h_ctr, sizeof h);
  if (h[AXR_CTR_BIN_SPILL] || h[AXR_CTR_LOST]) {
    bins_grow(b->bins, h[AXR_CTR_BIN_SPILL], h[AXR_CTR_LOST]);
    // a loss may come from the large-triangle list too: grow it as well
    if (h[AXR_CTR_LOST]) {
      const uint64_t want = std::max<uint64_t>(2ull * b->large_cap,
                                               (uint64_t)b->large_cap + h[AXR_CTR_LOST]);
      b->want_large = (uint32_t)std::min<uint64_t>(want, 1ull << 30);
    }
    HIPCHK(launch_fill_u32(b->d_counters, AXR_CTR_WORDS, 0, b->stream));
    std::memset(b->h_ctr, 0, sizeof h);
    if (h[AXR_CTR_LOST]) {
      b->incomplete = true;
      return fail(AXR_ECAP, "%u bin entries did not fit the spill list; it has been grown, "
                            "render the batch again",
                  h[AXR_CTR_LOST]);
    }
  }
  if (b->incomplete)
    return fail(AXR_ECAP, "the last render lost bin entries (the spill list has been grown): "
                          "render the batch again");
  return AXR_OK;
}

static int batch_out(axr_batch* b, int job, const void* src, size_t elem, void* dst) {
  if (!b || !dst) return fail(AXR_EINVAL, "batch readback: bad arguments");
  if (job < 0 || job >= b->njobs) return fail(AXR_EINVAL, "no job %d in the batch", job);
  int rc;
  if ((rc = axr_batch_wait(b, nullptr))) return rc;
  const JobInfo& in = b->info[(size_t)job];
  const size_t n = (size_t)in.W * in.H;
  HIPCHK(hipMemcpy(dst, (const char*)src + in.out_off * elem, n * elem, hipMemcpyDeviceToHost));
  return AXR_OK;
}

int axr_batch_get_rgba8(axr_batch* b, int job, uint32_t* out) {
  if (b && b->mode != SHADE_LDR8) return fail(AXR_ESTATE, "the batch was not rendered for RGBA8");
  return batch_out(b, job, b ? b->d_rgba : nullptr, 4, out);
}

int axr_batch_get_depth(axr_batch* b, int job, float* out) {
  return batch_out(b, job, b ? b->d_depth : nullptr, 4, out);
}

int axr_batch_get_ids(axr_batch* b, int job, uint32_t* out) {
  if (b && b->mode != SHADE_IDS) return fail(AXR_ESTATE, "the batch was not rendered for ids");
  return batch_out(b, job, b ? b->d_ids : nullptr, 4, out);
}

int axr_batch_get_hdr(axr_batch* b, int job, uint16_t* r, uint16_t* g, uint16_t* bl) {
  if (b && b->mode != SHADE_HDR) return fail(AXR_ESTATE, "the batch was not rendered for HDR");
  if (!b || !r || !g || !bl) return fail(AXR_EINVAL, "batch readback: bad arguments");
  int rc;
  if ((rc = batch_out(b, job, b->d_hdr, 2, r)) || (rc = batch_out(b, job, b->d_hdr + b->hdr_cap, 2, g)) ||
      (rc = batch_out(b, job, b->d_hdr + 2 * b->hdr_cap, 2, bl)))
    return rc;
  return AXR_OK;
}

int axr_render_cube(axr_batch* b, axr_ctx* scene, int slot, const float pos[3], int size,
                    int shadows, int pcf_taps) {
  if (!b || !scene || !pos) return fail(AXR_EINVAL, "renderCube: bad arguments");
  if (size < 16 || size > 8192 || (size % 16) != 0)
    return fail(AXR_EINVAL, "renderCube: face size %d must be a multiple of 16", size);
  int rc;
  if ((rc = check_ctx(scene))) return rc;
  if (slot >= (int)scene->refl.size()) return fail(AXR_EINVAL, "renderCube: no slot %d", slot);
  // Face f looks along s e_a (s = +1 for f < 3, -1 after, a = f mod 3) with
  // horizontal row -s e_u and vertical row s e_v, (u, v) = (z, y), (x, z),
  // (y, x) for a = x, y, z, at a 90 degree field of view: then texel (x, y)
  // of face f holds the direction the SSR shader's cube lookup maps there.
  axr_job jobs[6];
  for (int f = 0; f < 6; f++) {
    const float sgn = f < 3 ? 1.0f : -1.0f;
    const int a = f % 3;
    const int u = a == 0 ? 2 : a == 1 ? 0 : 1;
    const int v = a == 0 ? 1 : a == 1 ? 2 : 0;
    axr_job& j = jobs[f];
    std::memset(&j, 0, sizeof j);
    j.scene = scene;
    for (int k = 0; k < 3; k++) j.pos[k] = pos[k];
    for (int k = 0; k < 9; k++) j.vm[k] = 0.0f;
    j.vm[a] = sgn;
    j.vm[3 + u] = -sgn;
    j.vm[6 + v] = sgn;
    j.scale = (float)(size / 2);
    j.cax = j.cay = 3.0f;  // (W/2)/scale + 2
    j.width = j.height = size;
    j.shadows = shadows;
    j.pcf_taps = pcf_taps;
    j.exposure = 1.0f;
  }
  if ((rc = axr_batch_render(b, jobs, 6, AXR_OUT_HDR))) return rc;
  if (slot < 0) {
    if (scene->refl.size() >= 64) return fail(AXR_EINVAL, "renderCube: at most 64 cube maps");
    scene->refl.push_back(ReflTex{});
    slot = (int)scene->refl.size() - 1;
  }
  ReflTex& t = scene->refl[(size_t)slot];
  if (t.size != size) {
    HIPCHK(hipStreamSynchronize(scene->stream));  // an SSR draw may read the old strip
    dfree(t.d);
    if ((rc = dalloc(&t.d, (size_t)6 * size * size))) return rc;
    t.size = size;
  }
  HIPCHK(launch_pack_cube(b->d_hdr, b->d_hdr + b->hdr_cap, b->d_hdr + 2 * b->hdr_cap, size, t.d,
                          b->stream));
  // The scene's stream waits for this only where it reads the strip (an
  // SSR pass) or writes what the pass reads (wait_cube): its own drawAll
  // runs beside the cube pass.
  if (!scene->ev_cube)
    HIPCHK(hipEventCreateWithFlags(&scene->ev_cube, hipEventDisableTiming));
  HIPCHK(hipEventRecord(scene->ev_cube, b->stream));
  scene->cube_pending = true;
  HIPCHK(hipEventRecord(b->ev_done, b->stream));
  return slot;
}

int axr_batch_get_bytes(axr_batch* b, uint64_t* out) {
  if (!b || !out) return fail(AXR_EINVAL, "batch bytes: bad arguments");
  uint64_t n = 0;
  n += (uint64_t)b->rec_cap * sizeof(AxrTriRec);
  n += (uint64_t)b->px_cap * 12 + (uint64_t)b->hdr_cap * 6;
  n += (uint64_t)b->bins.ent_alloc * 4 + (uint64_t)b->bins.nb_alloc * 4 * 5 +
       (uint64_t)b->bins.B.spill_cap * 8;
  n += (uint64_t)b->jobs_cap * (sizeof(AxrJobDev) + 5 * sizeof(uint32_t));
  *out = n;
  return AXR_OK;
}

}  // extern "C"
