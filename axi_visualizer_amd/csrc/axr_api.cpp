// axr_api.cpp -- C ABI of libaxr.so (declared in include/axr.h).
//
// One context = one render target, its scene groups and shadow maps, and one
// HIP stream.  Per-frame work is enqueued without host synchronisation and
// without allocation: triangle counts and bin offsets stay on the device,
// group flags live in a device table rewritten only when they change.
//
// drawAll sets up and bins the triangles; the depth walk and shading run in
// one launch when colour is first needed (gamma or a readback), with gamma
// fused when gamma is what asked for it.  Shadow-map clears are applied
// lazily: the next shadow pass starts its bins from the clear depth instead
// of reading a cleared map.  Anything that would change the inputs of a
// draw that has not been shaded yet shades it first.
#include "axr_ctx.h"

namespace axri {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}


void drain_timing(axr_ctx* c) {
  if (c->trecs.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto& r : c->trecs) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto& acc = c->tacc[r.name];
      acc.first += ms;
      acc.second += 1;
    }
    c->free_events.push_back(r.a);
    c->free_events.push_back(r.b);
  }
  c->trecs.clear();
}

void bins_free(BinStore& S) {
  dfree(S.B.cnt);
  dfree(S.B.num);
  dfree(S.B.ent);
  dfree(S.B.order);
  dfree(S.B.spill);
  dfree(S.B.spill_ctr);
  S.nb_alloc = 0;
  S.ent_alloc = 0;
}

// Initial bucket capacity: `cap`, reduced to fit the budget for nb bins.
uint32_t fit_cap(uint32_t cap, size_t nb, int nseg) {
  const size_t per = kBucketBudget / (sizeof(uint32_t) * (size_t)nseg * (nb ? nb : 1));
  if (per < cap) cap = per < kMinCap ? kMinCap : (uint32_t)per;
  return cap;
}

// Makes S hold nb bins with the wanted capacities.  Only called when the
// bins of the previous pass of this kind are no longer needed: their
// entries are dropped if storage has to grow.  The stream is idle after.
int bins_prepare(BinStore& S, int nb, hipStream_t st) {
  const size_t slots = (size_t)S.nseg * nb;
  const size_t ents = (size_t)nb * S.want0 + (S.nseg > 1 ? (size_t)nb * S.want1 : 0);
  const bool grow_slots = nb > S.nb_alloc;
  const bool grow_ents = ents > S.ent_alloc || S.want0 != S.B.cap0 || S.want1 != S.B.cap1;
  const bool grow_spill = S.want_spill != S.B.spill_cap;
  if (!grow_slots && !grow_ents && !grow_spill && S.B.cnt) {
    S.B.nb = nb;
    return 0;
  }
  HIPCHK(hipStreamSynchronize(st));
  int rc;
  if (grow_slots || !S.B.cnt) {
    dfree(S.B.cnt);
    dfree(S.B.num);
    dfree(S.B.order);
    const size_t all = (size_t)S.nseg * nb;
    if ((rc = dalloc(&S.B.cnt, all)) || (rc = dalloc(&S.B.num, all)) ||
        (rc = dalloc(&S.B.order, (size_t)nb)))
      return rc;
    HIPCHK(launch_fill_u32(S.B.cnt, all, 0, st));
    HIPCHK(launch_fill_u32(S.B.num, all, 0, st));
    std::vector<uint32_t> h((size_t)nb);
    for (int i = 0; i < nb; i++) h[(size_t)i] = (uint32_t)i;
    HIPCHK(hipMemcpyAsync(S.B.order, h.data(), h.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice, st));
    S.nb_alloc = nb;
  }
  if (grow_ents || !S.B.ent) {
    dfree(S.B.ent);
    if ((rc = dalloc(&S.B.ent, ents))) return rc;
    S.ent_alloc = ents;
  }
  if (grow_spill || !S.B.spill) {
    dfree(S.B.spill);
    dfree(S.B.spill_ctr);
    if ((rc = dalloc(&S.B.spill, (size_t)S.want_spill)) || (rc = dalloc(&S.B.spill_ctr, 2)))
      return rc;
    HIPCHK(launch_fill_u32(S.B.spill_ctr, 2, 0, st));
  }
  S.B.cap0 = S.want0;
  S.B.cap1 = S.nseg > 1 ? S.want1 : 0;
  S.B.spill_cap = S.want_spill;
  S.B.nb = nb;
  (void)slots;
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

// Raises the wanted capacities after a pass of this kind spilled or lost
// entries (applied by the next bins_prepare).
void bins_grow(BinStore& S, uint32_t spilled, uint32_t lost) {
  if (spilled) {
    const size_t cur = (size_t)S.nb_alloc * (S.want0 + S.want1) * sizeof(uint32_t);
    if (2 * cur <= kBucketGrowLimit) {
      S.want0 *= 2;
      if (S.nseg > 1) S.want1 *= 2;
    }
  }
  if (lost) {
    uint64_t need = (uint64_t)S.want_spill + lost;
    uint64_t cap = 2 * (uint64_t)S.want_spill;
    while (cap < need) cap *= 2;
    S.want_spill = cap > 0x7FFFFFFFu ? 0x7FFFFFFFu : (uint32_t)cap;
  }
}

int check_ctx(axr_ctx* c) {
  if (!c) return fail(AXR_EINVAL, "null context");
  HIPCHK(hipSetDevice(c->device));
  return 0;
}

// Grows the concatenated vertex arrays and the per-triangle records.
int reserve_tris(axr_ctx* c, size_t extra) {
  const size_t need = (size_t)c->T + extra;
  if (need <= c->tri_cap) return 0;
  size_t cap = c->tri_cap ? c->tri_cap : 1024;
  while (cap < need) cap *= 2;
  float4 *np = nullptr, *na = nullptr;
  int rc;
  if ((rc = dalloc(&np, 3 * cap))) return rc;
  if ((rc = dalloc(&na, 6 * cap))) return rc;
  if (c->T) {
    HIPCHK(hipMemcpyAsync(np, c->d_vpos, (size_t)3 * c->T * sizeof(float4),
                          hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(na, c->d_vattr, (size_t)6 * c->T * sizeof(float4),
                          hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  dfree(c->d_vpos);
  dfree(c->d_vattr);
  c->d_vpos = np;
  c->d_vattr = na;
  c->tri_cap = cap;
  dfree(c->d_rec);
  dfree(c->d_cls);
  dfree(c->d_chunks);
  dfree(c->d_large);
  for (int k = 0; k < AXR_MAX_SHADOWS; k++) {
    dfree(c->d_shrec[k]);
    dfree(c->d_ok[k]);
  }
  c->last_ok = nullptr;
  if ((rc = dalloc(&c->d_rec, cap)) || (rc = dalloc(&c->d_cls, cap)) ||
      (rc = dalloc(&c->d_chunks, cap / 64)) || (rc = dalloc(&c->d_large, cap)))
    return rc;
  for (int k = 0; k < AXR_MAX_SHADOWS; k++)
    if ((rc = dalloc(&c->d_shrec[k], cap)) || (rc = dalloc(&c->d_ok[k], cap))) return rc;
  // cap is a power of two >= 1024: the byte arrays are whole words
  HIPCHK(launch_fill_u32((uint32_t*)c->d_cls, cap / 4, 0, c->stream));
  for (int k = 0; k < AXR_MAX_SHADOWS; k++)
    HIPCHK(launch_fill_u32((uint32_t*)c->d_ok[k], cap / 4, 0, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->records_valid = false;
  return 0;
}

AxrGroupDev group_dev(const HostGroup& g) {
  AxrGroupDev d{};
  d.base = g.base;
  d.ntri = g.ntri;
  d.flags = g.flags;
  d.tex_size = g.tex_size;
  d.mip_bias = g.mip_bias;
  d.spec_pow = g.spec_pow;
  d.tex = g.d_tex;
  return d;
}

// Orders the scene's stream after a cube pass still reading the scene.
int wait_cube(axr_ctx* c) {
  if (!c->cube_pending) return 0;
  HIPCHK(hipStreamWaitEvent(c->stream, c->ev_cube, 0));
  c->cube_pending = false;
  return 0;
}

int upload_groups(axr_ctx* c) {
  int rc;
  if ((rc = wait_cube(c))) return rc;
  std::vector<AxrGroupDev> h(c->groups.size());
  for (size_t i = 0; i < h.size(); i++) h[i] = group_dev(c->groups[i]);
  // Kernels in flight read the table: let them finish before rewriting it.
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!h.empty())
    HIPCHK(hipMemcpy(c->d_groups, h.data(), h.size() * sizeof(AxrGroupDev),
                     hipMemcpyHostToDevice));
  return 0;
}

// Per-bin depth walk + shading of the current draw into the requested target.
int bin_raster(axr_ctx* c, int mode, float exposure) {
  AxrDrawConst dc = c->last;
  const int empty = c->state == ST_CLEARED || c->state == ST_UNCLEARED;
  if (empty) {
    dc = AxrDrawConst{};
    dc.W = c->W;
    dc.H = c->H;
    dc.WC = c->WC;
    dc.HC = c->HC;
    dc.T = c->T;
    dc.ngroups = (uint32_t)c->groups.size();
    for (int k = 0; k < 3; k++) dc.clear_rgb[k] = c->clear_rgb[k];
  }
  Timed tm(c, mode == SHADE_LDR8 ? "raster_shade" : mode == SHADE_HDR ? "raster_hdr" : "raster_ids");
  AxrBins B = c->mbins.B;
  B.nb = c->NB;
  // Bins share blocks only when there are more of them than the device holds
  // blocks at once (two per CU); a frame of fewer bins is rastered a 16-px
  // quadrant per block (no tiers).
  const bool share = c->NB > AXR_QUAD_BINS_PER_CU * c->cu_count;
  // winner ids leave the depth buffer as it is: it may hold a later pass's
  // depths (SSR water) that the ids do not describe
  HIPCHK(launch_bin_raster(mode, c->d_rec, B, empty, c->d_vpos, c->d_vattr, c->d_groups, dc,
                           c->NBX, exposure, mode == SHADE_IDS ? nullptr : c->d_depth, c->d_r,
                           c->d_g, c->d_b, c->d_rgba, c->d_ids,
                           share ? c->d_large_ctl + 1 : nullptr, c->stream));
  return 0;
}

int materialize(axr_ctx* c, int mode, float exposure) {
  int rc;
  if (c->accum && c->state == ST_PENDING) {
    // a draw composed onto the frame: the pixels it wins, in HDR (with their
    // winner ids); gamma, if asked for, from the planes
    if ((rc = bin_raster(c, SHADE_ACCUM, 1.0f))) return rc;
    c->accum = false;
    c->state = ST_HDR;
    if (mode == SHADE_LDR8) {
      Timed tm(c, "gamma");
      HIPCHK(launch_gamma_pack(c->d_r, c->d_g, c->d_b, c->d_rgba, (size_t)c->W * c->H, exposure,
                               c->stream));
      c->state = ST_LDR8;
    }
    return 0;
  }
  if ((rc = bin_raster(c, mode, exposure))) return rc;
  c->state = mode == SHADE_LDR8 ? ST_LDR8 : ST_HDR;
  return 0;
}

// Before a drawAll onto a frame that already holds a draw (Ops.py:716-937
// draws into the current buffers; only clearZBuffer resets them): the frame
// so far gets its winner ids (the last draw's, from its bins) and its colour
// in the planes, so the new draw can be rastered onto depth, planes and ids.
int compose_prepare(axr_ctx* c) {
  int rc;
  if (!c->composed) {
    if (!c->records_valid)
      return fail(AXR_ESTATE, "drawAll onto a frame drawn before groups were added: clear first");
    if ((rc = bin_raster(c, SHADE_IDS, 1.0f))) return rc;
    c->composed = true;
  }
  if (c->state == ST_PENDING) {
    if ((rc = materialize(c, SHADE_HDR, 1.0f))) return rc;
  } else if (c->state == ST_LDR8) {
    HIPCHK(launch_expand_rgba8(c->d_rgba, c->d_r, c->d_g, c->d_b, (size_t)c->W * c->H,
                               c->stream));
    c->state = ST_LDR16;
  }
  return 0;
}

// Shades a pending draw before one of its inputs changes (the reference
// shades inside drawAll, so later calls cannot affect that frame).
int settle(axr_ctx* c) {
  if (c->state != ST_PENDING) return 0;
  return materialize(c, SHADE_HDR, 1.0f);
}

bool bound_to_pending(axr_ctx* c, int i) {
  if (c->state != ST_PENDING) return false;
  for (int k = 0; k < c->last.nshadow; k++)
    if (c->last.sh[k].map == c->sm[i].d_map) return true;
  return false;
}

// Writes the clear depth into a map whose clear is still pending.
int apply_shadow_clear(axr_ctx* c, int i) {
  HostShadow& s = c->sm[i];
  if (!s.clear_pending) return 0;
  int rc;
  if ((rc = wait_cube(c))) return rc;
  Timed tm(c, "shadow_clear");
  HIPCHK(launch_fill_u32(s.d_map, (size_t)s.dim * s.dim, ordf(AXR_DEPTH_CLEAR), c->stream));
  s.clear_pending = false;
  return 0;
}

// After a synchronising readback.  Spilled entries were still drawn (the
// buckets grow for later passes); entries lost because the spill list was
// full make the frame incomplete: report it and grow the list.  A loss is
// attributed to no pass, so the frame and every shadow map a pass wrote
// since the last check are marked incomplete: each stays so (its readbacks
// fail with AXR_ECAP) until it is drawn again from a clear.
// frame: the readback is of the frame (else of shadow map `map`).
int check_overflow(axr_ctx* c, bool frame, int map = -1) {
  uint32_t h[kCounters];
  HIPCHK(hipMemcpy(h, c->d_counters, sizeof h, hipMemcpyDeviceToHost));
  const uint32_t msp = h[AXR_CTR_BIN_SPILL], ssp = h[AXR_CTR_SH_SPILL], lost = h[AXR_CTR_LOST];
  if (msp || ssp || lost) {
    // A loss is not attributed to a kind: grow both spill lists.
    bins_grow(c->mbins, msp, lost);
    bins_grow(c->sbins, ssp, lost);
    const uint32_t zero[3] = {0, 0, 0};
    HIPCHK(hipMemcpy(c->d_counters + AXR_CTR_BIN_SPILL, zero, 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_counters + AXR_CTR_SH_SPILL, zero, 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_counters + AXR_CTR_LOST, zero, 4, hipMemcpyHostToDevice));
  }
  // (a frame drawn since the last clear may have lost them)
  if (lost && c->state != ST_UNCLEARED && c->state != ST_CLEARED) c->frame_incomplete = true;
  for (HostShadow& s : c->sm) {
    if (lost && s.unchecked) s.incomplete = true;
    s.unchecked = false;
  }
  if (lost)
    return fail(AXR_ECAP, "%u bin entries did not fit the spill list; it has been grown, redraw",
                lost);
  if (frame && c->frame_incomplete)
    return fail(AXR_ECAP, "the frame lost bin entries or samples an incomplete shadow map; "
                          "redraw it (clearZBuffer, drawAll)");
  if (map >= 0 && c->sm[map].incomplete)
    return fail(AXR_ECAP, "shadow map %d lost bin entries; redraw it (clearShadowMap, shadowMap)",
                map);
  return 0;
}

bool pass_pending(const axr_ctx* c, int i) {
  for (const auto& p : c->pend)
    if (p.map == i) return true;
  return false;
}

// Launches the pending shadow passes -- with the drawAll's setup when dc
// is given -- as one setup, one order and one shadow-depth launch.
// Draw constants of the context's current view and lights.
AxrDrawConst base_dc(const axr_ctx* c) {
  AxrDrawConst dc{};
  dc.view = c->view;
  dc.cull = make_cull(dc.view);
  for (int k = 0; k < 3; k++) {
    dc.lint[k] = c->lint[k];
    dc.ldir[k] = c->ldir[k];
    dc.clear_rgb[k] = c->clear_rgb[k];
  }
  dc.amb = c->amb;
  dc.W = c->W;
  dc.H = c->H;
  dc.WC = c->WC;
  dc.HC = c->HC;
  dc.T = c->T;
  dc.ngroups = (uint32_t)c->groups.size();
  dc.pcf_taps = 5;
  dc.flag_mask = 0xFFFFFFFFu;
  return dc;
}

int run_setup(axr_ctx* c, const AxrDrawConst* dc_main, const AxrGroupDev* table, uint32_t t0,
              uint32_t t1) {
  AxrPassSet ps{};
  int nb_total = 0;
  ps.n = (int)c->pend.size();
  for (int k = 0; k < ps.n; k++) {
    const PendingPass& q = c->pend[(size_t)k];
    const HostShadow& s = c->sm[q.map];
    AxrShadowPass& P = ps.p[k];
    P.cam = q.cam;
    P.cull = make_cull(P.cam);
    P.dim = s.dim;
    P.nbx = s.nbx;
    P.bias = q.bias;
    P.slot0 = (uint32_t)nb_total;
    P.from_map = q.from_map;
    P.cast = q.cast;
    P.rec = c->d_shrec[k];
    P.ok = c->d_ok[k];
    P.map = s.d_map;
    nb_total += s.nb;
    HostShadow& sw = c->sm[q.map];
    if (!q.from_map) sw.incomplete = false;
    sw.unchecked = true;
  }
  c->pend.clear();
  int rc;
  if (ps.n && (rc = wait_cube(c))) return rc;  // the passes write shadow maps
  if (ps.n && (rc = bins_prepare(c->sbins, nb_total, c->stream))) return rc;
  if (dc_main && (rc = bins_prepare(c->mbins, c->NB, c->stream))) return rc;
  AxrDrawConst dc{};
  if (dc_main) {
    dc = *dc_main;
  } else {
    dc.T = c->T;
    dc.ngroups = (uint32_t)c->groups.size();
  }
  {
    Timed tm(c, dc_main ? "setup_bin" : "shadow_setup_bin");
    // shadow passes take every caster: the range is the whole scene then
    if (ps.n) {
      t0 = 0;
      t1 = c->T;
    }
    HIPCHK(launch_setup(c->d_vpos, c->d_vattr, table ? table : c->d_groups, dc, dc_main ? 1 : 0,
                        c->d_rec, c->d_cls, c->mbins.B, c->NBX, ps, c->sbins.B, c->d_chunks, t0,
                        t1, c->d_large, c->d_large_ctl, c->stream));
  }
  {
    Timed tm(c, dc_main ? "bin_order" : "shadow_order");
    HIPCHK(launch_bin_order(dc_main ? &c->mbins.B : nullptr, ps.n ? &c->sbins.B : nullptr,
                            c->d_counters, c->d_large_ctl, c->d_large_ctl + 1, c->stream));
  }
  if (ps.n) {
    Timed tm(c, "shadow_raster");
    HIPCHK(launch_shadow_bins(c->sbins.B, ps, c->stream));
    c->last_ok = c->d_ok[ps.n - 1];
  }
  return 0;
}

int flush_shadows(axr_ctx* c) {
  if (c->pend.empty()) return 0;
  return run_setup(c, nullptr, nullptr, 0, c->T);
}

}  // namespace axri

extern "C" {

const char* axr_last_error(void) { return g_err.c_str(); }

int axr_create(int w, int h, axr_ctx** out) {
  if (!out) return fail(AXR_EINVAL, "null out pointer");
  *out = nullptr;
  if (w <= 0 || h <= 0)
    return fail(AXR_EINVAL, "width and height must be positive (%dx%d)", w, h);
  if ((int64_t)w * h > (int64_t)1 << 28) return fail(AXR_EINVAL, "frame too large");
  axr_ctx* c = new axr_ctx();
  HIPCHK(hipGetDevice(&c->device));
  HIPCHK(hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, c->device));
  c->W = w;
  c->H = h;
  c->WC = w / 16;
  c->HC = h / 16;
  c->NBX = (w + AXR_BIN - 1) / AXR_BIN;
  c->NB = c->NBX * ((h + AXR_BIN - 1) / AXR_BIN);
  const size_t np = (size_t)w * h;
  int rc = 0;
  auto bail = [&](int r) {
    axr_destroy(c);
    return r;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(AXR_EHIP, "hipStreamCreate failed"));
  if ((rc = dalloc(&c->d_depth, np)) || (rc = dalloc(&c->d_r, np)) ||
      (rc = dalloc(&c->d_g, np)) || (rc = dalloc(&c->d_b, np)) || (rc = dalloc(&c->d_rgba, np)) ||
      (rc = dalloc(&c->d_ids, np)) || (rc = dalloc(&c->d_counters, kCounters)) ||
      (rc = dalloc(&c->d_groups, AXR_MAX_GROUPS)) || (rc = dalloc(&c->d_large_ctl, 3)))
    return bail(rc);
  if (launch_fill_u32(c->d_counters, kCounters, 0, c->stream) != hipSuccess ||
      launch_fill_u32(c->d_large_ctl, 3, 0, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return bail(fail(AXR_EHIP, "initial fill failed"));
  c->mbins.nseg = 2;
  c->mbins.want0 = fit_cap(kMainCap0, (size_t)c->NB, 2);
  c->mbins.want1 = fit_cap(kMainCap1, (size_t)c->NB, 2);
  c->mbins.want_spill = kSpillCap;
  c->sbins.nseg = 1;
  c->sbins.want_spill = kSpillCap;
  if ((rc = bins_prepare(c->mbins, c->NB, c->stream))) return bail(rc);
  // Default view: identity basis at the origin, 90 degree fovx.
  const float vm[9] = {0, 0, 1, -1, 0, 0, 0, -1, 0};
  std::memcpy(c->view.vm, vm, sizeof vm);
  c->view.scale = (float)(w / 2);
  c->view.cax = 3.0f;
  c->view.cay = 3.0f;
  *out = c;
  return AXR_OK;
}

int axr_destroy(axr_ctx* c) {
  if (!c) return AXR_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) {
    (void)wait_cube(c);
    (void)hipStreamSynchronize(c->stream);
  }
  if (c->ev_cube) (void)hipEventDestroy(c->ev_cube);
  for (auto& r : c->trecs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : c->free_events) (void)hipEventDestroy(e);
  for (auto& g : c->groups) dfree(g.d_tex);
  for (auto& s : c->sm) dfree(s.d_map);
  dfree(c->d_vpos);
  dfree(c->d_vattr);
  dfree(c->d_groups);
  dfree(c->d_rec);
  dfree(c->d_cls);
  dfree(c->d_chunks);
  for (int k = 0; k < AXR_MAX_SHADOWS; k++) {
    dfree(c->d_shrec[k]);
    dfree(c->d_ok[k]);
  }
  bins_free(c->mbins);
  bins_free(c->sbins);
  dfree(c->d_depth);
  dfree(c->d_r);
  dfree(c->d_g);
  dfree(c->d_b);
  dfree(c->d_rgba);
  dfree(c->d_ids);
  dfree(c->d_counters);
  dfree(c->d_large_ctl);
  dfree(c->d_large);
  dfree(c->d_light);
  if (c->d_post) (void)hipFree(c->d_post);
  if (c->d_snap) (void)hipFree(c->d_snap);
  dfree(c->d_groups_ssr);
  for (ReflTex& t : c->refl) dfree(t.d);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return AXR_OK;
}

int axr_set_scale_cull(axr_ctx* c, float s, float cax, float cay) {
  if (!c) return fail(AXR_EINVAL, "null context");
  c->view.scale = s;
  c->view.cax = cax;
  c->view.cay = cay;
  return AXR_OK;
}

int axr_add_texture_group(axr_ctx* c, const float* xyz, const float* uv, const float* vn,
                          int ntri, const uint16_t* r, const uint16_t* g, const uint16_t* b,
                          int64_t ntexels, int tex_size, int is_mip, int* tn_out) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (!xyz || !uv || !vn || !r || !g || !b || ntri < 0 || ntexels <= 0)
    return fail(AXR_EINVAL, "addTextureGroup: bad arguments");
  if ((int)c->groups.size() >= AXR_MAX_GROUPS)
    return fail(AXR_ECAP, "at most %d groups per context", AXR_MAX_GROUPS);
  if ((uint64_t)c->T + (uint64_t)ntri >= (1ull << 30))
    return fail(AXR_ECAP, "at most 2^30 triangles per context");
  if ((rc = wait_cube(c))) return rc;  // geometry and textures may be reallocated
  // Texture shape: N*N texels, the full smallest-first chain of N*N, or
  // its N*N base level (chain built here).
  if (is_mip < 0 || is_mip > 2) return fail(AXR_EINVAL, "is_mip must be 0, 1 or 2");
  int64_t chain = ntexels;
  if (is_mip) {
    // Trilinear reads levels mip and mip+1 <= texSize, so a chain needs >= 2 levels.
    if (tex_size < 1 || tex_size > 14) return fail(AXR_EINVAL, "mip texSize out of range");
    chain = (((int64_t)1 << (2 * (tex_size + 1))) - 1) / 3;
    const int64_t want = is_mip == 1 ? chain : ((int64_t)1 << (2 * tex_size));
    if (ntexels != want)
      return fail(AXR_EINVAL, "%s of 2^%d needs %lld texels, got %lld",
                  is_mip == 1 ? "mip chain" : "base level", tex_size, (long long)want,
                  (long long)ntexels);
  } else {
    if (tex_size <= 0 || (tex_size & (tex_size - 1)) ||
        (int64_t)tex_size * tex_size != ntexels)
      return fail(AXR_EINVAL, "texture must be N*N with N a power of two");
  }
  if ((rc = settle(c))) return rc;
  // pending shadow passes cast the groups that existed when they were made
  if ((rc = flush_shadows(c))) return rc;
  if ((rc = reserve_tris(c, (size_t)ntri))) return rc;
  const size_t nv = (size_t)3 * ntri;
  std::vector<float4> hp(nv), ha(2 * nv);
  for (size_t v = 0; v < nv; v++) {
    hp[v] = make_float4(xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2], 0.0f);
    ha[2 * v] = make_float4(vn[3 * v], vn[3 * v + 1], vn[3 * v + 2], uv[2 * v]);
    ha[2 * v + 1] = make_float4(uv[2 * v + 1], 1.0f, 1.0f, 1.0f);
  }
  if (nv) {
    HIPCHK(hipMemcpy(c->d_vpos + (size_t)3 * c->T, hp.data(), nv * sizeof(float4),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_vattr + (size_t)6 * c->T, ha.data(), 2 * nv * sizeof(float4),
                     hipMemcpyHostToDevice));
  }
  std::vector<ushort4> ht((size_t)ntexels);
  for (int64_t i = 0; i < ntexels; i++) ht[i] = make_ushort4(r[i], g[i], b[i], 0);
  HostGroup hg;
  hg.base = c->T;
  hg.ntri = (uint32_t)ntri;
  hg.tex_size = tex_size;
  hg.is_mip = is_mip ? 1 : 0;
  hg.ntexels = chain;
  if ((rc = dalloc(&hg.d_tex, (size_t)chain))) return rc;
  // the given texels are the chain, or its last (largest) level
  HIPCHK(hipMemcpy(hg.d_tex + (chain - ntexels), ht.data(), ht.size() * sizeof(ushort4),
                   hipMemcpyHostToDevice));
  if (is_mip == 2) {
    HIPCHK(launch_create_mips(hg.d_tex, tex_size + 1, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->groups.push_back(hg);
  c->T += (uint32_t)ntri;
  HIPCHK(launch_chunk_bounds(c->d_vpos, c->T, c->d_chunks, c->stream));
  if ((rc = upload_groups(c))) return rc;
  if (tn_out) *tn_out = (int)c->groups.size() - 1;
  return AXR_OK;
}

int axr_set_group_material(axr_ctx* c, int tn, float mip_bias, int spec_pow) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (tn < 0 || tn >= (int)c->groups.size()) return fail(AXR_EINVAL, "no group %d", tn);
  if (spec_pow < 0) return fail(AXR_EINVAL, "spec_pow must be >= 0");
  if ((rc = settle(c))) return rc;
  c->groups[tn].mip_bias = mip_bias;
  c->groups[tn].spec_pow = spec_pow;
  return upload_groups(c);
}

int axr_set_primary_light(axr_ctx* c, const float I[3], const float D[3]) {
  if (!c || !I || !D) return fail(AXR_EINVAL, "setPrimaryLight: bad arguments");
  for (int k = 0; k < 3; k++) {
    c->lint[k] = I[k];
    c->ldir[k] = D[k];
  }
  return AXR_OK;
}

int axr_vert_light(axr_ctx* c, const uint8_t* mask, int ngroups, const float* dirI,
                   const float* dirD, int ndir, const float* pointI, const float* pointP,
                   int npoint, const float* spotI, const float* spotD, const float* spotP,
                   int nspot, float amb) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (ngroups != (int)c->groups.size() || !mask)
    return fail(AXR_EINVAL, "vertLight: mask must cover all %d groups", (int)c->groups.size());
  if (ndir < 0 || npoint < 0 || nspot < 0) return fail(AXR_EINVAL, "negative light count");
  if ((ndir && (!dirI || !dirD)) || (npoint && (!pointI || !pointP)) ||
      (nspot && (!spotI || !spotD || !spotP)))
    return fail(AXR_EINVAL, "vertLight: missing light arrays");
  if ((rc = settle(c))) return rc;
  std::vector<float> h;
  auto put = [&](const float* p, int n) {
    size_t off = h.size();
    for (int i = 0; i < 3 * n; i++) h.push_back(p[i]);
    return off;
  };
  const size_t oDI = put(dirI, ndir), oDD = put(dirD, ndir);
  const size_t oPI = put(pointI, npoint), oPP = put(pointP, npoint);
  const size_t oSI = put(spotI, nspot), oSD = put(spotD, nspot), oSP = put(spotP, nspot);
  if (h.size() > c->light_cap) {
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(c->d_light);
    if ((rc = dalloc(&c->d_light, h.size()))) return rc;
    c->light_cap = h.size();
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!h.empty())
    HIPCHK(hipMemcpy(c->d_light, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  AxrLights L;
  L.ndir = ndir;
  L.npoint = npoint;
  L.nspot = nspot;
  L.amb = amb;
  L.dirI = c->d_light + oDI;
  L.dirD = c->d_light + oDD;
  L.pointI = c->d_light + oPI;
  L.pointP = c->d_light + oPP;
  L.spotI = c->d_light + oSI;
  L.spotD = c->d_light + oSD;
  L.spotP = c->d_light + oSP;
  c->amb = amb;
  if ((rc = wait_cube(c))) return rc;  // the lights are written into the vertex attributes
  for (int g = 0; g < ngroups; g++) {
    if (!mask[g]) continue;
    const HostGroup& hg = c->groups[g];
    HIPCHK(launch_vertlight(c->d_vpos, c->d_vattr, 3 * hg.base, 3 * hg.ntri, L, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return AXR_OK;
}

int axr_set_pos(axr_ctx* c, const float pos[3]) {
  if (!c || !pos) return fail(AXR_EINVAL, "setPos: bad arguments");
  for (int k = 0; k < 3; k++) c->view.pos[k] = pos[k];
  return AXR_OK;
}

int axr_set_vm(axr_ctx* c, const float vm[9]) {
  if (!c || !vm) return fail(AXR_EINVAL, "setVM: bad arguments");
  for (int k = 0; k < 9; k++) c->view.vm[k] = vm[k];
  return AXR_OK;
}

int axr_set_clear_color(axr_ctx* c, const uint16_t rgb[3]) {
  if (!c || !rgb) return fail(AXR_EINVAL, "setClearColor: bad arguments");
  for (int k = 0; k < 3; k++) c->clear_rgb[k] = rgb[k];
  return AXR_OK;
}

int axr_clear_zbuffer(axr_ctx* c) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  // The bins of the next drawAll start empty (the order pass zeroes them),
  // and the per-bin walk starts from the clear key: nothing to launch.
  c->state = ST_CLEARED;
  c->frame_incomplete = false;
  c->composed = c->accum = false;
  return AXR_OK;
}

int axr_draw_all(axr_ctx* c, const uint32_t* shader_flags, const uint8_t* mask, int ngroups,
                 const int* shadow_ids, int nshadow, int pcf_taps) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (ngroups != (int)c->groups.size())
    return fail(AXR_EINVAL, "drawAll: %d shader entries for %d groups", ngroups,
                (int)c->groups.size());
  if (!shader_flags) return fail(AXR_EINVAL, "drawAll: null shaders");
  if (c->state == ST_UNCLEARED) return fail(AXR_ESTATE, "drawAll needs a clearZBuffer first");
  if (nshadow < 0 || nshadow > AXR_SHADOW_BIND)
    return fail(AXR_EINVAL, "drawAll: at most %d shadow maps", AXR_SHADOW_BIND);
  if (pcf_taps != 5 && pcf_taps != 9) return fail(AXR_EINVAL, "pcf_taps must be 5 or 9");
  // a draw onto the frame so far (no clearZBuffer since the last drawAll)
  const bool onto = c->state != ST_CLEARED;
  if (onto && (rc = compose_prepare(c))) return rc;
  AxrDrawConst dc = base_dc(c);
  dc.pcf_taps = pcf_taps;
  bool any_shadow = false, changed = false;
  std::vector<uint32_t> nf((size_t)ngroups);
  for (int g = 0; g < ngroups; g++) {
    const uint32_t f = shader_flags[g] & (AXR_SH_MIP | AXR_SH_SHADOW | AXR_SH_PHONG | AXR_SH_CULL);
    if ((f & AXR_SH_MIP) && !c->groups[g].is_mip)
      return fail(AXR_EINVAL, "group %d: mip shader on a texture without mips", g);
    if ((f & AXR_SH_SHADOW) && !(mask && mask[g])) any_shadow = true;
    nf[g] = f | ((mask && mask[g]) ? 0u : AXR_GRP_DRAW);
    changed |= nf[g] != c->groups[g].flags;
  }
  dc.nshadow = 0;
  if (any_shadow) {
    for (int k = 0; k < nshadow; k++) {
      const int i = shadow_ids ? shadow_ids[k] : k;
      if (i < 0 || i >= AXR_MAX_SHADOWS || !c->sm[i].exists || !c->sm[i].placed)
        return fail(AXR_ESTATE, "drawAll: shadow map %d missing or not placed", i);
      if ((rc = apply_shadow_clear(c, i))) return rc;
      if (c->sm[i].incomplete) c->frame_incomplete = true;
      AxrShadowDev& sd = dc.sh[dc.nshadow++];
      sd.map = c->sm[i].d_map;
      sd.dim = c->sm[i].dim;
      sd.wS = c->sm[i].dim / 2;
      sd.cam = c->sm[i].cam;
    }
  }
  if (changed) {
    for (int g = 0; g < ngroups; g++) c->groups[g].flags = nf[g];
    if ((rc = upload_groups(c))) return rc;
  }
  // the draw's setup runs with the pending shadow passes (which also
  // write any map this draw samples before its shading runs)
  if (c->T > 0 && (rc = run_setup(c, &dc, nullptr, 0, c->T))) return rc;
  c->last = dc;
  c->records_valid = true;
  if (c->T > 0) {
    c->state = ST_PENDING;
    c->accum = onto;
  } else if (!onto) {
    c->state = ST_CLEARED;
  }
  return AXR_OK;
}

int axr_set_refl_tex(axr_ctx* c, int slot, const uint16_t* r, const uint16_t* g,
                     const uint16_t* b, int size) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (!r || !g || !b) return fail(AXR_EINVAL, "setReflTex: null plane");
  if (size < 2 || size > 8192 || (size & 1))
    return fail(AXR_EINVAL, "setReflTex: face size %d must be even, 2..8192", size);
  if (slot >= (int)c->refl.size()) return fail(AXR_EINVAL, "setReflTex: no slot %d", slot);
  if (slot < 0) {
    if (c->refl.size() >= 64) return fail(AXR_EINVAL, "setReflTex: at most 64 cube maps");
    c->refl.push_back(ReflTex{});
    slot = (int)c->refl.size() - 1;
  }
  ReflTex& t = c->refl[(size_t)slot];
  const size_t n = (size_t)6 * size * size;
  std::vector<ushort4> h(n);
  for (size_t i = 0; i < n; i++) h[i] = make_ushort4(r[i], g[i], b[i], 0);
  if ((rc = wait_cube(c))) return rc;  // a cube pass may still write the strip
  HIPCHK(hipStreamSynchronize(c->stream));
  if (t.size != size) {
    dfree(t.d);
    if ((rc = dalloc(&t.d, n))) return rc;
    t.size = size;
  }
  HIPCHK(hipMemcpy(t.d, h.data(), n * sizeof(ushort4), hipMemcpyHostToDevice));
  return slot;
}

int axr_get_refl_tex(axr_ctx* c, int slot, uint16_t* r, uint16_t* g, uint16_t* b) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (slot < 0 || slot >= (int)c->refl.size() || !c->refl[(size_t)slot].d)
    return fail(AXR_EINVAL, "getReflTex: no cube map in slot %d", slot);
  const ReflTex& t = c->refl[(size_t)slot];
  const size_t n = (size_t)6 * t.size * t.size;
  std::vector<ushort4> h(n);
  if ((rc = wait_cube(c))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(h.data(), t.d, n * sizeof(ushort4), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; i++) {
    if (r) r[i] = h[i].x;
    if (g) g[i] = h[i].y;
    if (b) b[i] = h[i].z;
  }
  return t.size;
}

int axr_draw_ssr(axr_ctx* c, const int* groups, const int* refl, int n, int steps, int step,
                 float absorb, int fresnel) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (n < 1 || !groups || !refl) return fail(AXR_EINVAL, "drawSSR: no groups");
  if (steps < 1 || steps > (1 << 20) || step < 1 || step > steps)
    return fail(AXR_EINVAL, "drawSSR: steps %d / step %d", steps, step);
  if (!(absorb == absorb)) return fail(AXR_EINVAL, "drawSSR: absorb is NaN");
  const int ng = (int)c->groups.size();
  std::vector<int> slot_of((size_t)ng, -1);
  for (int k = 0; k < n; k++) {
    const int g = groups[k], s = refl[k];
    if (g < 0 || g >= ng) return fail(AXR_EINVAL, "drawSSR: no group %d", g);
    if (c->groups[(size_t)g].is_mip)
      return fail(AXR_EINVAL, "drawSSR: group %d: the water texture must not be a mip chain", g);
    if (s < 0 || s >= (int)c->refl.size() || !c->refl[(size_t)s].d)
      return fail(AXR_EINVAL, "drawSSR: group %d: no cube map in slot %d", g, s);
    slot_of[(size_t)g] = s;
  }
  if (c->state == ST_UNCLEARED) return fail(AXR_ESTATE, "drawSSR before any clearZBuffer");
  if (c->state == ST_LDR8 || c->state == ST_LDR16)
    return fail(AXR_ESTATE, "drawSSR after gamma: the pass works on the HDR frame");
  if ((rc = flush_shadows(c))) return rc;
  if (c->state == ST_CLEARED || c->state == ST_PENDING) {
    if ((rc = materialize(c, SHADE_HDR, 1.0f))) return rc;
  }
  // the water samples cube strips a cube pass may still be rendering (the
  // opaque frame above did not need to wait for it)
  if ((rc = wait_cube(c))) return rc;
  const size_t np = (size_t)c->W * c->H;
  // planes of the snapshot start 16-byte aligned: a stride of np rounded up to 8
  const size_t np8 = (np + 7) & ~(size_t)7;
  const size_t snap = np8 * (sizeof(float) + 3 * sizeof(uint16_t));
  if (c->snap_bytes < snap) {
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->d_snap) (void)hipFree(c->d_snap);
    c->d_snap = nullptr;
    c->snap_bytes = 0;
    HIPCHK(hipMalloc(&c->d_snap, snap));
    c->snap_bytes = snap;
  }
  AxrSsrConst S{};
  S.F0 = (const float*)c->d_snap;
  S.R0 = (const uint16_t*)((char*)c->d_snap + np8 * sizeof(float));
  S.G0 = S.R0 + np8;
  S.B0 = S.G0 + np8;
  S.steps = steps;
  S.step = step;
  S.absorb = absorb;
  S.fresnel = fresnel ? 1 : 0;
  {
    Timed tm(c, "ssr_copy");
    // one launch: 16-byte vectors, the last np % 8 pixels one by one
    HIPCHK(launch_snap_frame(c->d_depth, c->d_r, c->d_g, c->d_b, np, (float*)S.F0,
                             (uint16_t*)S.R0, (uint16_t*)S.G0, (uint16_t*)S.B0, c->stream));
  }
  // Only the water groups take part in this draw: its own group table (the
  // drawAll's table stays as it is) and the water groups' triangle range.
  std::vector<AxrGroupDev> tab((size_t)ng);
  uint32_t t0 = c->T, t1 = 0;
  for (int g = 0; g < ng; g++) {
    const HostGroup& G = c->groups[(size_t)g];
    const int s = slot_of[(size_t)g];
    tab[(size_t)g] = group_dev(G);
    tab[(size_t)g].flags = s >= 0 ? AXR_GRP_DRAW : 0u;
    if (s >= 0) {
      tab[(size_t)g].refl = c->refl[(size_t)s].d;
      tab[(size_t)g].refl_dim = c->refl[(size_t)s].size / 2;
      t0 = std::min(t0, G.base);
      t1 = std::max(t1, G.base + G.ntri);
    }
  }
  if (!c->d_groups_ssr || c->h_groups_ssr.size() != tab.size() ||
      std::memcmp(c->h_groups_ssr.data(), tab.data(), tab.size() * sizeof(AxrGroupDev)) != 0) {
    HIPCHK(hipStreamSynchronize(c->stream));  // an SSR raster may read the old table
    if (c->h_groups_ssr.size() != tab.size()) {
      dfree(c->d_groups_ssr);
      if ((rc = dalloc(&c->d_groups_ssr, tab.size()))) return rc;
    }
    HIPCHK(hipMemcpy(c->d_groups_ssr, tab.data(), tab.size() * sizeof(AxrGroupDev),
                     hipMemcpyHostToDevice));
    c->h_groups_ssr = tab;
  }
  AxrDrawConst dc = base_dc(c);
  if (t1 > t0) {
    if ((rc = run_setup(c, &dc, c->d_groups_ssr, t0, t1))) return rc;
    Timed tm(c, "ssr_raster");
    AxrBins B = c->mbins.B;
    B.nb = c->NB;
    HIPCHK(launch_ssr(c->d_rec, B, c->d_vpos, c->d_vattr, c->d_groups_ssr, dc, S, c->NBX,
                      c->d_depth, c->d_r, c->d_g, c->d_b, c->stream));
  }
  c->state = ST_HDR;
  c->records_valid = false;
  return AXR_OK;
}

int axr_add_shadow_map(axr_ctx* c, int i, int size, float scale, float amb) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (i < 0 || i >= AXR_MAX_SHADOWS) return fail(AXR_EINVAL, "shadow map index %d", i);
  if (size < 8 || size > 16384 || (size & 1)) return fail(AXR_EINVAL, "shadow map size %d", size);
  if (bound_to_pending(c, i) && (rc = settle(c))) return rc;
  if (pass_pending(c, i) && (rc = flush_shadows(c))) return rc;
  HostShadow& s = c->sm[i];
  HIPCHK(hipStreamSynchronize(c->stream));
  dfree(s.d_map);
  if ((rc = dalloc(&s.d_map, (size_t)size * size))) return rc;
  s.exists = true;
  s.placed = false;
  s.clear_pending = true;  // addShadowMap fills 256*256-1 (Ops.py:1049)
  s.dim = size;
  s.nbx = (size + AXR_SHADOW_BIN - 1) / AXR_SHADOW_BIN;
  s.nb = s.nbx * s.nbx;
  s.scale = scale;
  s.cam.scale = scale;
  if (amb >= 0) c->amb = amb;
  // Shadow buckets are sized for every map being in one batch of passes.
  int nb_all = 0;
  for (const auto& m : c->sm)
    if (m.exists) nb_all += m.nb;
  const uint32_t cap = fit_cap(kShadowCap, (size_t)nb_all, 1);
  if (!c->sbins.want0 || cap < c->sbins.want0) c->sbins.want0 = cap;
  HIPCHK(hipStreamSynchronize(c->stream));
  return AXR_OK;
}

int axr_place_shadow_map(axr_ctx* c, int i, const float pos[3], const float vm[9], float amb) {
  if (!c || !pos || !vm) return fail(AXR_EINVAL, "placeShadowMap: bad arguments");
  if (i < 0 || i >= AXR_MAX_SHADOWS || !c->sm[i].exists)
    return fail(AXR_EINVAL, "no shadow map %d", i);
  HostShadow& s = c->sm[i];
  for (int k = 0; k < 3; k++) s.cam.pos[k] = pos[k];
  for (int k = 0; k < 9; k++) s.cam.vm[k] = vm[k];
  s.cam.scale = s.scale;
  s.placed = true;
  if (amb >= 0) c->amb = amb;
  return AXR_OK;
}

int axr_clear_shadow_map(axr_ctx* c, int i) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (i < 0 || i >= AXR_MAX_SHADOWS || !c->sm[i].exists)
    return fail(AXR_EINVAL, "no shadow map %d", i);
  if (bound_to_pending(c, i) && (rc = settle(c))) return rc;
  if (pass_pending(c, i) && (rc = flush_shadows(c))) return rc;
  c->sm[i].clear_pending = true;
  c->sm[i].incomplete = false;
  return AXR_OK;
}

int axr_shadow_map(axr_ctx* c, int i, const uint8_t* cast, int ngroups, float bias) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (i < 0 || i >= AXR_MAX_SHADOWS || !c->sm[i].exists || !c->sm[i].placed)
    return fail(AXR_EINVAL, "shadow map %d missing or not placed", i);
  if (ngroups != (int)c->groups.size() || !cast)
    return fail(AXR_EINVAL, "shadowMap: whichCast must cover all groups");
  if (bound_to_pending(c, i) && (rc = settle(c))) return rc;
  // a second pass on a map starts from the first one's result
  if (pass_pending(c, i) && (rc = flush_shadows(c))) return rc;
  AxrCastMask m{};
  for (int g = 0; g < ngroups; g++)
    if (cast[g]) m.bits[g >> 5] |= 1u << (g & 31);
  if (c->T == 0) return AXR_OK;
  HostShadow& s = c->sm[i];
  PendingPass q;
  q.map = i;
  q.cast = m;
  q.bias = bias;
  q.cam = s.cam;
  q.from_map = s.clear_pending ? 0 : 1;
  c->pend.push_back(q);
  s.clear_pending = false;
  return AXR_OK;
}

int axr_gamma(axr_ctx* c, float ex) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  switch (c->state) {
    case ST_CLEARED:
    case ST_PENDING:
      return materialize(c, SHADE_LDR8, ex);
    case ST_HDR: {
      // gamma'd colour straight into RGBA8 (getFrame expands it back)
      Timed tm(c, "gamma");
      HIPCHK(launch_gamma_pack(c->d_r, c->d_g, c->d_b, c->d_rgba, (size_t)c->W * c->H, ex,
                               c->stream));
      c->state = ST_LDR8;
      return AXR_OK;
    }
    case ST_LDR16: {
      Timed tm(c, "gamma");
      HIPCHK(launch_gamma_planes(c->d_r, c->d_g, c->d_b, (size_t)c->W * c->H, ex, c->stream));
      c->state = ST_LDR16;
      return AXR_OK;
    }
    case ST_LDR8: {
      Timed tm(c, "gamma");
      const size_t n = (size_t)c->W * c->H;
      HIPCHK(launch_expand_rgba8(c->d_rgba, c->d_r, c->d_g, c->d_b, n, c->stream));
      HIPCHK(launch_gamma_planes(c->d_r, c->d_g, c->d_b, n, ex, c->stream));
      c->state = ST_LDR16;
      return AXR_OK;
    }
    default:
      return fail(AXR_ESTATE, "gamma before any clearZBuffer");
  }
}

// The planes a post pass works on: HDR colour, or gamma'd colour after
// gamma() (the reference's post kernels read whatever RO/GO/BO hold).
static int post_planes(axr_ctx* c, const char* what) {
  int rc;
  if (c->state == ST_UNCLEARED) return fail(AXR_ESTATE, "%s before any clearZBuffer", what);
  if ((rc = flush_shadows(c))) return rc;
  if (c->state == ST_CLEARED || c->state == ST_PENDING) return materialize(c, SHADE_HDR, 1.0f);
  if (c->state == ST_LDR8) {
    HIPCHK(launch_expand_rgba8(c->d_rgba, c->d_r, c->d_g, c->d_b, (size_t)c->W * c->H,
                               c->stream));
    c->state = ST_LDR16;
  }
  return 0;
}

static int post_scratch(axr_ctx* c, size_t bytes) {
  if (c->post_bytes >= bytes) return 0;
  if (c->d_post) {
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipFree(c->d_post);
    c->d_post = nullptr;
    c->post_bytes = 0;
  }
  HIPCHK(hipMalloc(&c->d_post, bytes));
  c->post_bytes = bytes;
  return 0;
}

int axr_blur(axr_ctx* c) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (c->W < 4 || c->H < 8 || (c->W & 1) || (c->H & 1))
    return fail(AXR_EINVAL, "blur needs an even frame of at least 4x8, got %dx%d", c->W, c->H);
  if ((rc = post_planes(c, "blur"))) return rc;
  if ((rc = post_scratch(c, 6 * sizeof(uint16_t) * (size_t)(c->W / 2) * (c->H / 2)))) return rc;
  Timed tm(c, "blur");
  HIPCHK(launch_blur(c->d_r, c->d_g, c->d_b, c->W, c->H, (uint16_t*)c->d_post, c->stream));
  return AXR_OK;
}

int axr_bloom(axr_ctx* c, int levels, float threshold, float white, float strength) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (levels < 1 || levels > 12)
    return fail(AXR_EINVAL, "bloom: levels must be 1..12, got %d", levels);
  if ((c->W % (1 << levels)) || (c->H % (1 << levels)))
    return fail(AXR_EINVAL, "bloom: %dx%d is not a multiple of 2^%d", c->W, c->H, levels);
  if (!(white > 0.0f) || !(strength >= 0.0f) || threshold != threshold)
    return fail(AXR_EINVAL, "bloom: white must be > 0 and strength >= 0");
  if ((rc = post_planes(c, "bloom"))) return rc;
  if ((rc = post_scratch(c, 2 * bloom_halves(c->W, c->H, levels)))) return rc;
  Timed tm(c, "bloom");
  HIPCHK(launch_bloom(c->d_r, c->d_g, c->d_b, c->W, c->H, levels, threshold, white, strength,
                      c->d_post, c->stream));
  return AXR_OK;
}

int axr_get_frame(axr_ctx* c, uint16_t* r, uint16_t* g, uint16_t* b, float* depth) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  const size_t n = (size_t)c->W * c->H;
  if (c->state == ST_UNCLEARED) return fail(AXR_ESTATE, "getFrame before any clearZBuffer");
  if ((rc = flush_shadows(c))) return rc;
  if (c->state == ST_CLEARED || c->state == ST_PENDING) {
    if ((rc = materialize(c, SHADE_HDR, 1.0f))) return rc;
  } else if (c->state == ST_LDR8) {
    HIPCHK(launch_expand_rgba8(c->d_rgba, c->d_r, c->d_g, c->d_b, n, c->stream));
    c->state = ST_LDR16;
  }
  if (r) HIPCHK(hipMemcpyAsync(r, c->d_r, n * 2, hipMemcpyDeviceToHost, c->stream));
  if (g) HIPCHK(hipMemcpyAsync(g, c->d_g, n * 2, hipMemcpyDeviceToHost, c->stream));
  if (b) HIPCHK(hipMemcpyAsync(b, c->d_b, n * 2, hipMemcpyDeviceToHost, c->stream));
  if (depth) HIPCHK(hipMemcpyAsync(depth, c->d_depth, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return check_overflow(c, true);
}

int axr_get_rgba8(axr_ctx* c, uint32_t* rgba) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (!rgba) return fail(AXR_EINVAL, "null output");
  if ((rc = flush_shadows(c))) return rc;
  const size_t n = (size_t)c->W * c->H;
  if (c->state == ST_LDR8) {
    HIPCHK(hipMemcpyAsync(rgba, c->d_rgba, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return check_overflow(c, true);
  }
  if (c->state == ST_LDR16) {
    // the planes hold the gamma'd colour (post passes after gamma): pack
    // them on the device, so only the packed frame crosses the link
    HIPCHK(launch_pack_rgba8(c->d_r, c->d_g, c->d_b, c->d_rgba, n, c->stream));
    HIPCHK(hipMemcpyAsync(rgba, c->d_rgba, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return check_overflow(c, true);
  }
  return fail(AXR_ESTATE, "getRGBA8 needs gamma() first");
}

int axr_get_ids(axr_ctx* c, uint32_t* ids) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (!ids) return fail(AXR_EINVAL, "null output");
  if (c->state == ST_UNCLEARED) return fail(AXR_ESTATE, "getIds before any clearZBuffer");
  if (!c->composed && c->state != ST_CLEARED && !c->records_valid)
    return fail(AXR_ESTATE, "getIds: groups were added since the last drawAll");
  if ((rc = flush_shadows(c))) return rc;
  const size_t n = (size_t)c->W * c->H;
  if (c->composed) {
    // several draws since the clear: their winners are kept as each is composed
    if (c->state == ST_PENDING && (rc = materialize(c, SHADE_HDR, 1.0f))) return rc;
  } else if ((rc = bin_raster(c, SHADE_IDS, 1.0f))) {  // the current draw's winners
    return rc;
  }
  HIPCHK(hipMemcpyAsync(ids, c->d_ids, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return check_overflow(c, true);
}

int axr_get_shadow_map(axr_ctx* c, int i, float* out) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (i < 0 || i >= AXR_MAX_SHADOWS || !c->sm[i].exists || !out)
    return fail(AXR_EINVAL, "no shadow map %d", i);
  if ((rc = flush_shadows(c))) return rc;
  if ((rc = apply_shadow_clear(c, i))) return rc;
  const size_t n = (size_t)c->sm[i].dim * c->sm[i].dim;
  float* d = nullptr;
  if ((rc = dalloc(&d, n))) return rc;
  hipError_t e = launch_decode_shadow(c->sm[i].d_map, d, n, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, n * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dfree(d);
  if (e != hipSuccess) return fail(AXR_EHIP, "getSHM: %s", hipGetErrorString(e));
  return check_overflow(c, false, i);
}

int axr_host_alloc(uint64_t bytes, void** out) {
  if (!out) return fail(AXR_EINVAL, "null out pointer");
  *out = nullptr;
  if (bytes == 0) return fail(AXR_EINVAL, "host_alloc: zero bytes");
  HIPCHK(hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault));
  return AXR_OK;
}

int axr_host_free(void* p) {
  if (p) HIPCHK(hipHostFree(p));
  return AXR_OK;
}

int axr_sync(axr_ctx* c) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if ((rc = flush_shadows(c))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return AXR_OK;
}

int axr_get_stats(axr_ctx* c, uint32_t out[6]) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (!out) return fail(AXR_EINVAL, "null output");
  if ((rc = flush_shadows(c))) return rc;
  if (c->T) HIPCHK(launch_stats(c->d_cls, c->last_ok, c->T, c->d_counters + AXR_CTR_STATS, c->stream));
  uint32_t h[kCounters];
  HIPCHK(hipMemcpyAsync(h, c->d_counters, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  out[0] = c->T ? h[AXR_CTR_STATS] : 0;
  out[1] = c->T ? h[AXR_CTR_STATS + 1] : 0;
  out[2] = c->T ? h[AXR_CTR_STATS + 2] : 0;
  out[3] = h[AXR_CTR_BIN_TOTAL];
  out[4] = h[AXR_CTR_LOST];
  out[5] = h[AXR_CTR_SH_TOTAL];
  return AXR_OK;
}

int axr_get_texels(axr_ctx* c, int tn, uint16_t* r, uint16_t* g, uint16_t* b) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (tn < 0 || tn >= (int)c->groups.size() || !r || !g || !b)
    return fail(AXR_EINVAL, "getTexels: bad arguments");
  const HostGroup& hg = c->groups[(size_t)tn];
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<ushort4> h((size_t)hg.ntexels);
  HIPCHK(hipMemcpy(h.data(), hg.d_tex, h.size() * sizeof(ushort4), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < h.size(); i++) {
    r[i] = h[i].x;
    g[i] = h[i].y;
    b[i] = h[i].z;
  }
  return AXR_OK;
}

int axr_get_vertex_light(axr_ctx* c, int tn, float* out) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (tn < 0 || tn >= (int)c->groups.size() || !out)
    return fail(AXR_EINVAL, "getVertLight: bad arguments");
  const HostGroup& hg = c->groups[(size_t)tn];
  const size_t nv = 3 * (size_t)hg.ntri;
  HIPCHK(hipStreamSynchronize(c->stream));
  // vattr[2v + 1] = (v, Ix, Iy, Iz)
  std::vector<float4> h(2 * nv);
  if (nv)
    HIPCHK(hipMemcpy(h.data(), c->d_vattr + 2 * 3 * (size_t)hg.base, h.size() * sizeof(float4),
                     hipMemcpyDeviceToHost));
  for (size_t v = 0; v < nv; v++) {
    out[3 * v] = h[2 * v + 1].y;
    out[3 * v + 1] = h[2 * v + 1].z;
    out[3 * v + 2] = h[2 * v + 1].w;
  }
  return AXR_OK;
}

int axr_set_shaders(axr_ctx* c, const uint32_t* shader_flags, const uint8_t* mask,
                    int ngroups) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (ngroups != (int)c->groups.size() || !shader_flags)
    return fail(AXR_EINVAL, "setShaders: %d entries for %d groups", ngroups,
                (int)c->groups.size());
  bool changed = false;
  std::vector<uint32_t> nf((size_t)ngroups);
  for (int g = 0; g < ngroups; g++) {
    const uint32_t f = shader_flags[g] & (AXR_SH_MIP | AXR_SH_SHADOW | AXR_SH_PHONG | AXR_SH_CULL);
    if ((f & AXR_SH_MIP) && !c->groups[g].is_mip)
      return fail(AXR_EINVAL, "group %d: mip shader on a texture without mips", g);
    nf[g] = f | ((mask && mask[g]) ? 0u : AXR_GRP_DRAW);
    changed |= nf[g] != c->groups[g].flags;
  }
  if (!changed) return AXR_OK;
  if ((rc = settle(c))) return rc;
  for (int g = 0; g < ngroups; g++) c->groups[g].flags = nf[g];
  return upload_groups(c);
}

int axr_get_bytes(axr_ctx* c, uint64_t* out) {
  if (!c || !out) return fail(AXR_EINVAL, "bytes: bad arguments");
  uint64_t n = 0;
  n += (uint64_t)c->tri_cap * (3 * 16 + 6 * 16 + sizeof(AxrTriRec) + 1);
  n += (uint64_t)c->tri_cap * AXR_MAX_SHADOWS * (sizeof(AxrShRec) + 1);
  for (const auto& g : c->groups) n += (uint64_t)g.ntexels * 8;
  for (const auto& s : c->sm)
    if (s.exists) n += (uint64_t)s.dim * s.dim * 4;
  for (const BinStore* S : {&c->mbins, &c->sbins})
    n += (uint64_t)S->ent_alloc * 4 + (uint64_t)S->nb_alloc * 4 * 5 + (uint64_t)S->B.spill_cap * 8;
  n += (uint64_t)c->W * c->H * (4 + 2 * 3 + 4 + 4);
  *out = n;
  return AXR_OK;
}

int axr_set_bin_capacity(axr_ctx* c, uint32_t small, uint32_t large, uint32_t shadow,
                         uint32_t spill) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  if (small) c->mbins.want0 = small;
  if (large) c->mbins.want1 = large;
  if (shadow) c->sbins.want0 = shadow;
  if (spill) {
    c->mbins.want_spill = spill;
    c->sbins.want_spill = spill;
  }
  return AXR_OK;
}

int axr_set_timing(axr_ctx* c, int enable) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  drain_timing(c);
  c->tacc.clear();
  c->timing = enable != 0;
  return AXR_OK;
}

int axr_get_timing(axr_ctx* c, char* names, int name_len, double* total_ms, int64_t* count,
                   int n) {
  int rc;
  if ((rc = check_ctx(c))) return rc;
  drain_timing(c);
  int k = 0;
  for (auto& kv : c->tacc) {
    if (k >= n) break;
    if (names && name_len > 0) {
      std::strncpy(names + (size_t)k * name_len, kv.first.c_str(), name_len - 1);
      names[(size_t)k * name_len + name_len - 1] = 0;
    }
    if (total_ms) total_ms[k] = kv.second.first;
    if (count) count[k] = kv.second.second;
    k++;
  }
  return k;
}

}  // extern "C"
