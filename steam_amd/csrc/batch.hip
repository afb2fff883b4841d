// Batched flux advance: many independent instances (each with its own plan,
// kernel spectrum, workspace, flux and keyed noise) advanced together.
//
// Every pass runs as one launch per kernel variant: the instances whose
// geometry selects the same compiled kernel (register z plan, register or
// generic x/y FFT) share a launch over the concatenation of their
// workgroups; a workgroup finds its instance by a binary search over the
// group's first-workgroup offsets.  Descriptors live in device memory and
// are refreshed from a pinned host copy at the start of every advance, so
// the whole advance can be captured once into a hipGraph and replayed with
// new noise keys.
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "conv_plan.h"

using namespace steam;

namespace {

// BP_NOISE_EXT: the noise fields of the instances created with their own
// noise buffer, generated by steam_batch_noise (not part of an advance)
enum BatchPass { BP_NOISE = 0, BP_Z, BP_Y, BP_X, BP_ZINV, BP_FIN, BP_POST, BP_NOISE_EXT,
                 BP_COUNT };

struct Group {
  int offset;   // first descriptor of the group in the flat array
  int count;    // instances
  int wgs;      // total workgroups of the launch
  int variant;  // kernel selector (pass-specific)
  int arg;      // second selector (radix set, post rescale flag)
  size_t lds;   // dynamic LDS of the generic kernels (group maximum)
  int threads;  // block size of the register x pass
};

// one instance's choices, fixed at creation
struct ItemPlan {
  const steam_conv_plan* p;
  bool zreg, tile, gen_noise, post, post_rescale, xf, apply;
  int wg_noise_ext;  // workgroups of its strided noise launch (steam_batch_noise)
};

}  // namespace

struct steam_batch {
  int n = 0;
  std::vector<PassDesc> items;   // per-instance descriptors (host master copy)
  std::vector<ItemPlan> plans;
  std::vector<int> slot_item;    // flat descriptor slot -> instance
  std::vector<std::vector<int>> item_slots;  // instance -> its flat slots
  std::vector<Group> groups[BP_COUNT];
  PassDesc* host = nullptr;      // pinned staging of the flat array
  PassDesc* dev = nullptr;
  int nslots = 0;
  hipEvent_t uploaded = nullptr;
  bool pending_upload = false;
  // a pass whose instances fall into several kernel variants runs them on
  // parallel branches: the largest group on the caller's stream, the others
  // on these side streams, forked from and joined back into it
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t fork = nullptr;
  hipEvent_t join[2] = {nullptr, nullptr};
};

namespace {

int wg_count(const ItemPlan& ip, int pass) {
  const ConvGeom& g = ip.p->g;
  switch (pass) {
    case BP_NOISE: return ip.p->wg_noise;
    case BP_NOISE_EXT: return ip.wg_noise_ext;
    case BP_Z:
    case BP_ZINV: return ip.zreg ? g.wg_zr : g.wg_z;
    case BP_Y: return g.wg_y;
    case BP_X: return g.wg_x;
    case BP_FIN: return 1;
    default: return ip.p->wg_post;
  }
}

// pass slot used by the kernels' descriptor lookup
int desc_pass(int pass) {
  switch (pass) {
    case BP_NOISE:
    case BP_NOISE_EXT: return PASS_NOISE;
    case BP_Z: return PASS_Z;
    case BP_Y: return PASS_Y;
    case BP_X: return PASS_X;
    case BP_ZINV: return PASS_ZINV;
    case BP_FIN: return PASS_FIN;
    default: return PASS_POST;
  }
}

// (variant, arg, threads) of an instance in a pass; variant < 0: not run
std::tuple<int, int, int> variant_of(const ItemPlan& ip, int pass) {
  const ConvGeom& g = ip.p->g;
  switch (pass) {
    case BP_NOISE: return {ip.gen_noise ? 0 : -1, 0, 256};
    case BP_NOISE_EXT: return {ip.gen_noise ? -1 : 0, 0, 256};
    case BP_Z:
    case BP_ZINV:
      // plan length, odd nz, the served flux options (unit entering flux,
      // stored apply rescale)
      if (ip.zreg) return {1, 4 * g.n2 + 2 * (g.nz & 1) + (ip.xf ? 1 : 0), ZR_NTH};
      return {ip.tile ? 2 : 3, radix_set(g.az), Z_NTH};
    case BP_Y:
      if (g.reg_y) return {1, g.reg_y, 256};
      if (g.mr_y) return {3, g.mr_y, xyplan_threads(g.mr_y)};
      return {2, radix_set(g.ay), generic_xy_nth(radix_set(g.ay))};
    case BP_X:
      if (g.reg_x) return {1, g.reg_x, (g.nx / 8 / xreg_lanes(g.reg_x)) * g.Tr};
      if (g.mr_x) return {3, g.mr_x, xyplan_threads(g.mr_x)};
      return {2, radix_set(g.ax), generic_xy_nth(radix_set(g.ax))};
    case BP_FIN: return {ip.apply ? 1 : 0, 0, ip.apply ? FIN_APPLY_NTH : FIN_NTH};
    default: return {ip.post ? 0 : -1, ip.post_rescale ? 1 : 0, FFT_NTH};
  }
}

size_t lds_of(const ItemPlan& ip, int pass) {
  switch (pass) {
    case BP_Z: return ip.zreg ? 0 : ip.p->lds_z;
    case BP_ZINV: return ip.zreg ? 0 : ip.p->lds_zinv;
    case BP_Y: return (ip.p->g.reg_y || ip.p->g.mr_y) ? 0 : ip.p->lds_y;
    case BP_X: return (ip.p->g.reg_x || ip.p->g.mr_x) ? 0 : ip.p->lds_x;
    default: return 0;
  }
}

// Workgroup cap of a batch's noise generated ahead: STEAM_BATCH_NOISE_WGS
// from the environment when set (read once), else 1024 (four per CU; the
// 48-member batch of 146M cells over 40 pipelined steps: 2.06-2.09 ms at
// 1024, 2.14-2.15 at 768, 2.14-2.17 at 1280, 2.19-2.20 at 1536, 2.21-2.25
// at 2048, 2.70 at 512 and 4.0 at 256, where the noise falls behind)
static int batch_noise_wgs_cap() {
  static const int cap = [] {
    const char* e = getenv("STEAM_BATCH_NOISE_WGS");
    return e && *e ? atoi(e) : 1024;
  }();
  return cap;
}

void build_groups(steam_batch* b) {
  std::vector<PassDesc> flat;
  b->slot_item.clear();
  // the noise generated ahead runs beside the transforms of the step
  // before: its launch gets at most the cap's workgroups, shared out by
  // instance size, each instance's workgroups striding over its lines
  {
    int64_t total = 0;
    for (int i = 0; i < b->n; ++i)
      if (!b->plans[i].gen_noise) total += b->plans[i].p->wg_noise;
    const int cap = batch_noise_wgs_cap();
    for (int i = 0; i < b->n; ++i) {
      ItemPlan& ip = b->plans[i];
      const int full = ip.p->wg_noise;
      int n = full;
      if (cap > 0 && total > cap)
        n = (int)std::max<int64_t>(1, std::min<int64_t>(full, (int64_t)cap * full / total));
      ip.wg_noise_ext = n;
      b->items[i].wg_noise_n = n;
    }
  }
  for (int pass = 0; pass < BP_COUNT; ++pass) {
    std::map<std::tuple<int, int, int>, std::vector<int>> by;
    for (int i = 0; i < b->n; ++i) {
      auto v = variant_of(b->plans[i], pass);
      if (std::get<0>(v) >= 0) by[v].push_back(i);
    }
    for (auto& kv : by) {
      Group gr;
      gr.offset = (int)flat.size();
      gr.count = (int)kv.second.size();
      gr.variant = std::get<0>(kv.first);
      gr.arg = std::get<1>(kv.first);
      gr.threads = std::get<2>(kv.first);
      gr.lds = 0;
      int base = 0;
      for (int i : kv.second) {
        PassDesc d = b->items[i];
        d.wg_base[desc_pass(pass)] = base;
        base += wg_count(b->plans[i], pass);
        gr.lds = std::max(gr.lds, lds_of(b->plans[i], pass));
        flat.push_back(d);
        b->slot_item.push_back(i);
      }
      gr.wgs = base;
      b->groups[pass].push_back(gr);
    }
  }
  b->nslots = (int)flat.size();
  memcpy(b->host, flat.data(), flat.size() * sizeof(PassDesc));
  b->item_slots.assign(b->n, {});
  for (int s = 0; s < b->nslots; ++s) b->item_slots[b->slot_item[s]].push_back(s);
}

// refresh instance i's slots of the flat array from its master copy (keeps
// each slot's workgroup offsets)
void refresh_item(steam_batch* b, int i) {
  for (int s : b->item_slots[i]) {
    int wb[7];
    memcpy(wb, b->host[s].wg_base, sizeof(wb));
    b->host[s] = b->items[i];
    memcpy(b->host[s].wg_base, wb, sizeof(wb));
  }
}

int launch_group(steam_batch* b, int pass, const Group& gr, hipStream_t s, int inv) {
  {
    BatchSrc src{b->dev + gr.offset, gr.count};
    const dim3 grid(gr.wgs);
    switch (pass) {
      case BP_NOISE:
        hipLaunchKernelGGL((noise_field_kernel<BatchSrc>), grid, dim3(256), 0, s, src);
        break;
      case BP_NOISE_EXT:
        hipLaunchKernelGGL((noise_field_kernel<BatchSrc, true>), grid, dim3(256), 0, s, src);
        break;
      case BP_Z:
        if (gr.variant == 1) {
          with_zplan(gr.arg >> 2, [&](auto zp) {
            using ZP = decltype(zp);
            switch (gr.arg & 3) {
              case 0: hipLaunchKernelGGL((zfwd_reg_kernel<BatchSrc, MODE_FLUX, ZP, false>), grid,
                                         dim3(ZR_NTH), 0, s, src); break;
              case 1: hipLaunchKernelGGL((zfwd_reg_kernel<BatchSrc, MODE_FLUX, ZP, false, true>),
                                         grid, dim3(ZR_NTH), 0, s, src); break;
              case 2: hipLaunchKernelGGL((zfwd_reg_kernel<BatchSrc, MODE_FLUX, ZP, true>), grid,
                                         dim3(ZR_NTH), 0, s, src); break;
              default: hipLaunchKernelGGL((zfwd_reg_kernel<BatchSrc, MODE_FLUX, ZP, true, true>),
                                          grid, dim3(ZR_NTH), 0, s, src); break;
            }
          });
        } else if (gr.variant == 2) {
          STEAM_RS_DISPATCH(gr.arg, hipLaunchKernelGGL((zfwd_tile_kernel<BatchSrc, MODE_FLUX, RS_>),
                                                       grid, dim3(Z_NTH), gr.lds, s, src));
        } else {
          STEAM_RS_DISPATCH(gr.arg, hipLaunchKernelGGL((zfwd_kernel<BatchSrc, MODE_FLUX, RS_>),
                                                       grid, dim3(Z_NTH), gr.lds, s, src));
        }
        break;
      case BP_Y:
        if (gr.variant == 1) {
          STEAM_LOGN_DISPATCH(gr.arg, hipLaunchKernelGGL((ypass_reg_kernel<BatchSrc, LN_>), grid,
                                                         dim3(256), 0, s, src, inv));
        } else if (gr.variant == 3) {
          with_xyplan(gr.arg, [&](auto xp) {
            hipLaunchKernelGGL((ypass_mr_kernel<BatchSrc, decltype(xp)>), grid,
                               dim3(xy_nth<decltype(xp)>()), 0, s, src, inv);
          });
        } else {
          STEAM_RS_DISPATCH_XY(gr.arg, hipLaunchKernelGGL((ypass_kernel<BatchSrc, RS_>), grid,
                                                          dim3(generic_xy_nth(RS_)), gr.lds, s,
                                                          src, inv));
        }
        break;
      case BP_X:
        if (gr.variant == 1) {
          STEAM_LOGN_DISPATCH(gr.arg, hipLaunchKernelGGL(
              (xpass_reg_kernel<BatchSrc, 0, LN_, xreg_lanes(LN_)>), grid, dim3(gr.threads), 0, s,
              src));
        } else if (gr.variant == 3) {
          with_xyplan(gr.arg, [&](auto xp) {
            hipLaunchKernelGGL((xpass_mr_kernel<BatchSrc, 0, decltype(xp)>), grid,
                               dim3(xy_nth<decltype(xp)>()), 0, s, src);
          });
        } else {
          STEAM_RS_DISPATCH_XY(gr.arg, hipLaunchKernelGGL((xpass_kernel<BatchSrc, 0, RS_>), grid,
                                                          dim3(generic_xy_nth(RS_)), gr.lds, s,
                                                          src));
        }
        break;
      case BP_ZINV:
        if (gr.variant == 1) {
          with_zplan(gr.arg >> 2, [&](auto zp) {
            using ZP = decltype(zp);
            switch (gr.arg & 3) {
              case 0: hipLaunchKernelGGL((zinv_reg_kernel<BatchSrc, MODE_FLUX, ZP, false>), grid,
                                         dim3(ZR_NTH), 0, s, src); break;
              case 1: hipLaunchKernelGGL((zinv_reg_kernel<BatchSrc, MODE_FLUX, ZP, false, true>),
                                         grid, dim3(ZR_NTH), 0, s, src); break;
              case 2: hipLaunchKernelGGL((zinv_reg_kernel<BatchSrc, MODE_FLUX, ZP, true>), grid,
                                         dim3(ZR_NTH), 0, s, src); break;
              default: hipLaunchKernelGGL((zinv_reg_kernel<BatchSrc, MODE_FLUX, ZP, true, true>),
                                          grid, dim3(ZR_NTH), 0, s, src); break;
            }
          });
        } else if (gr.variant == 2) {
          STEAM_RS_DISPATCH(gr.arg, hipLaunchKernelGGL((zinv_tile_kernel<BatchSrc, MODE_FLUX, RS_>),
                                                       grid, dim3(Z_NTH), gr.lds, s, src));
        } else {
          STEAM_RS_DISPATCH(gr.arg, hipLaunchKernelGGL((zinv_kernel<BatchSrc, MODE_FLUX, RS_>),
                                                       grid, dim3(Z_NTH), gr.lds, s, src));
        }
        break;
      case BP_FIN:
        if (gr.variant == 1)
          hipLaunchKernelGGL((finalize_apply_kernel<BatchSrc>), grid, dim3(FIN_APPLY_NTH), 0, s,
                             src);
        else
          hipLaunchKernelGGL((finalize_kernel<BatchSrc>), grid, dim3(FIN_NTH), 0, s, src);
        break;
      default:
        hipLaunchKernelGGL((post_kernel<BatchSrc>), grid, dim3(FFT_NTH), 0, s, src, gr.arg);
        break;
    }
    STEAM_LAUNCH_CHECK("batched pass");
  }
  return STEAM_OK;
}

// One pass: its variant groups are independent (different instances), so
// with several groups the largest runs on s and the others on the side
// streams, forked from s before and joined into it after (a captured graph
// records them as parallel branches).  Short launches of small members then
// overlap the large ones instead of running one after another.
int launch_pass(steam_batch* b, int pass, hipStream_t s, int inv) {
  const std::vector<Group>& gs = b->groups[pass];
  if (gs.size() < 2 || !b->side[0]) {
    for (const Group& gr : gs) {
      int rc = launch_group(b, pass, gr, s, inv);
      if (rc) return rc;
    }
    return STEAM_OK;
  }
  size_t big = 0;
  for (size_t i = 1; i < gs.size(); ++i)
    if (gs[i].wgs > gs[big].wgs) big = i;
  STEAM_HIP(hipEventRecord(b->fork, s));
  bool used[2] = {false, false};
  int k = 0;
  for (size_t i = 0; i < gs.size(); ++i) {
    if (i == big) continue;
    const int j = k++ % 2;
    if (!used[j]) {
      STEAM_HIP(hipStreamWaitEvent(b->side[j], b->fork, 0));
      used[j] = true;
    }
    int rc = launch_group(b, pass, gs[i], b->side[j], inv);
    if (rc) return rc;
  }
  int rc = launch_group(b, pass, gs[big], s, inv);
  if (rc) return rc;
  for (int j = 0; j < 2; ++j) {
    if (!used[j]) continue;
    STEAM_HIP(hipEventRecord(b->join[j], b->side[j]));
    STEAM_HIP(hipStreamWaitEvent(s, b->join[j], 0));
  }
  return STEAM_OK;
}

int make_item(const steam_batch_item& it, PassDesc& d, ItemPlan& ip) {
  const steam_conv_plan* p = it.plan;
  if (!p || !it.spectrum_dev || !it.workspace_dev || !it.flux_dev)
    return set_error(STEAM_EINVAL, "steam_batch_create: NULL plan, spectrum, workspace or flux");
  int rc = check_noise_params(&it.noise);
  if (rc) return rc;
  d = base_desc(p, it.workspace_dev);
  d.mode = MODE_FLUX;
  d.K = (const float2*)((const char*)it.spectrum_dev + KSPEC_DATA_OFF);
  d.kfmt = (const int*)it.spectrum_dev;
  d.ktile = p->ktile;
  d.out = it.flux_dev;
  d.sigma_in = it.sigma_in;
  d.pend = (it.flags & STEAM_CHAIN_PENDING)
               ? (const DevLedger*)((const char*)it.workspace_dev + p->led_off)
               : nullptr;
  d.noise = make_noise_dev(&it.noise);
  d.sk = it.sk_dev;
  d.apply = it.ledger != nullptr;
  if (d.apply) d.led_in = ledger_to_dev(it.ledger);
  d.unit_in = (it.flags & STEAM_UNIT_FLUX) != 0;
  if (d.unit_in && (it.flags & STEAM_CHAIN_PENDING))
    return set_error(STEAM_EINVAL, "steam_batch_create: a unit entering flux has no pending "
                                   "rescale to chain");
  d.clip_win = (it.flags & STEAM_CLIP_IN_WINDOW) != 0;
  const int64_t* w = it.window;
  if (w[0] || w[1] || w[2] || w[3]) {
    if (w[0] < 0 || w[1] > p->nx || w[0] >= w[1] || w[2] < 0 || w[3] > p->ny || w[2] >= w[3])
      return set_error(STEAM_EINVAL, "steam_batch_create: window out of range");
    for (int i = 0; i < 4; ++i) d.win[i] = (int)w[i];
  }
  if (it.noise_dev) d.nfield = const_cast<float*>(it.noise_dev);
  auto a8 = [](const void* q) { return ((uintptr_t)q & 7) == 0; };
  auto a16 = [](const void* q) { return ((uintptr_t)q & 15) == 0; };
  ip.p = p;
  ip.zreg = p->g.zreg && ((p->g.nz & 1) || (a8(d.out) && a8(d.nfield) && a8(d.sk)));
  ip.tile = !ip.zreg && p->g.tile_z && a16(d.out) && a16(d.nfield) && a16(d.sk);
  ip.gen_noise = it.noise_dev == nullptr;
  const bool defer = (it.flags & STEAM_DEFER_RESCALE) != 0;
  // a served tile (unit entering flux, apply mode): the inverse z pass
  // stores the rescaled flux (other apply-mode items keep the post pass)
  if (d.apply && !defer && d.unit_in) set_out_scale(d, d.led_in);
  ip.post = (!defer && !d.out_scale) || it.sk_dev != nullptr;
  ip.post_rescale = !defer && !d.out_scale;
  ip.xf = d.unit_in || d.out_scale;
  ip.apply = d.apply;
  d.npart = ip.zreg ? p->g.wg_zr : p->g.wg_z;
  return STEAM_OK;
}

}  // namespace

extern "C" int steam_batch_create(const steam_batch_item* items, int n,
                                  steam_batch** out) {
  if (!out) return set_error(STEAM_EINVAL, "steam_batch_create: NULL output");
  *out = nullptr;
  if (!items || n < 1) return set_error(STEAM_EINVAL, "steam_batch_create: no items");
  steam_batch* b = new steam_batch();
  b->n = n;
  b->items.resize(n);
  b->plans.resize(n);
  for (int i = 0; i < n; ++i) {
    int rc = make_item(items[i], b->items[i], b->plans[i]);
    if (rc) {
      delete b;
      return rc;
    }
  }
  // upper bound on descriptor slots: every pass holds each instance once
  const size_t bytes = (size_t)BP_COUNT * n * sizeof(PassDesc);
  hipError_t e = hipHostMalloc((void**)&b->host, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    delete b;
    return check_hip(e, "hipHostMalloc(batch descriptors)");
  }
  e = hipMalloc((void**)&b->dev, bytes);
  if (e != hipSuccess) {
    (void)hipHostFree(b->host);
    delete b;
    return check_hip(e, "hipMalloc(batch descriptors)");
  }
  e = hipEventCreateWithFlags(&b->uploaded, hipEventDisableTiming);
  if (e != hipSuccess) {
    (void)hipFree(b->dev);
    (void)hipHostFree(b->host);
    delete b;
    return check_hip(e, "hipEventCreate(batch)");
  }
  build_groups(b);
  // side streams (default priority) and the fork/join events, only for a
  // batch with a pass of several kernel variants
  bool several = false;
  for (int p = 0; p < BP_COUNT; ++p) several = several || b->groups[p].size() > 1;
  if (several) {
    e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (int j = 0; j < 2 && e == hipSuccess; ++j) {
      e = hipStreamCreateWithFlags(&b->side[j], hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[j], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
      const int rc = check_hip(e, "hipStreamCreate(batch side streams)");
      steam_batch_destroy(b);
      return rc;
    }
  }
  *out = b;
  return STEAM_OK;
}

extern "C" int steam_batch_destroy(steam_batch* b) {
  if (!b) return STEAM_OK;
  if (b->pending_upload) (void)hipEventSynchronize(b->uploaded);
  if (b->uploaded) (void)hipEventDestroy(b->uploaded);
  for (int j = 0; j < 2; ++j) {
    if (b->side[j]) {
      (void)hipStreamSynchronize(b->side[j]);
      (void)hipStreamDestroy(b->side[j]);
    }
    if (b->join[j]) (void)hipEventDestroy(b->join[j]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->dev) (void)hipFree(b->dev);
  if (b->host) (void)hipHostFree(b->host);
  delete b;
  return STEAM_OK;
}

extern "C" int steam_batch_set_noise(steam_batch* b, int i,
                                     const steam_noise_params* noise) {
  if (!b || i < 0 || i >= b->n)
    return set_error(STEAM_EINVAL, "steam_batch_set_noise: bad batch or index");
  int rc = check_noise_params(noise);
  if (rc) return rc;
  // the previous advance's upload must have read the staging copy
  if (b->pending_upload) {
    STEAM_HIP(hipEventSynchronize(b->uploaded));
    b->pending_upload = false;
  }
  b->items[i].noise = make_noise_dev(noise);
  refresh_item(b, i);
  return STEAM_OK;
}

extern "C" int steam_batch_set_crop(steam_batch* b, int i, float* out_dev,
                                    const int64_t* origin, const int64_t* shape) {
  if (!b || i < 0 || i >= b->n)
    return set_error(STEAM_EINVAL, "steam_batch_set_crop: bad batch or index");
  PassDesc& d = b->items[i];
  if (out_dev) {
    if (!origin || !shape) return set_error(STEAM_EINVAL, "steam_batch_set_crop: NULL box");
    if (!d.out_scale || d.sk || !b->plans[i].zreg)
      return set_error(STEAM_EINVAL,
                       "steam_batch_set_crop: only an apply-mode instance with a unit entering "
                       "flux, no S_k and a register z plan stores a crop");
    const int64_t dims[3] = {d.g.nx, d.g.ny, d.g.nz};
    for (int a = 0; a < 3; ++a)
      if (origin[a] < 0 || shape[a] < 1 || origin[a] + shape[a] > dims[a])
        return set_error(STEAM_EINVAL, "steam_batch_set_crop: box outside the grid");
  }
  // the previous advance's upload must have read the staging copy
  if (b->pending_upload) {
    STEAM_HIP(hipEventSynchronize(b->uploaded));
    b->pending_upload = false;
  }
  d.crop_out = out_dev;
  for (int a = 0; a < 3; ++a) {
    d.crop_o[a] = out_dev ? (int)origin[a] : 0;
    d.crop_n[a] = out_dev ? (int)shape[a] : 0;
  }
  refresh_item(b, i);
  return STEAM_OK;
}

extern "C" int steam_batch_mark_replay(steam_batch* b, steam_stream_t stream) {
  if (!b) return set_error(STEAM_EINVAL, "steam_batch_mark_replay: NULL batch");
  // a replay's captured upload reads the staging copy when the replay runs:
  // the next steam_batch_set_noise waits for this point of the stream
  STEAM_HIP(hipEventRecord(b->uploaded, (hipStream_t)stream));
  b->pending_upload = true;
  return STEAM_OK;
}

extern "C" int steam_batch_info(const steam_batch* b, int* launches,
                                int64_t* descriptor_bytes) {
  if (!b) return set_error(STEAM_EINVAL, "steam_batch_info: NULL batch");
  int n = 1;  // descriptor upload
  for (int p = 0; p < BP_COUNT; ++p)
    if (p != BP_NOISE_EXT) n += (int)b->groups[p].size() * (p == BP_Y ? 2 : 1);
  if (launches) *launches = n;
  if (descriptor_bytes) *descriptor_bytes = (int64_t)b->nslots * (int64_t)sizeof(PassDesc);
  return STEAM_OK;
}

extern "C" int steam_batch_noise(steam_batch* b, steam_stream_t stream) {
  if (!b) return set_error(STEAM_EINVAL, "steam_batch_noise: NULL batch");
  if (b->groups[BP_NOISE_EXT].empty())
    return set_error(STEAM_EINVAL, "steam_batch_noise: no instance has its own noise buffer");
  hipStream_t s = (hipStream_t)stream;
  STEAM_HIP(hipMemcpyAsync(b->dev, b->host, (size_t)b->nslots * sizeof(PassDesc),
                           hipMemcpyHostToDevice, s));
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  STEAM_HIP(hipStreamIsCapturing(s, &cs));
  if (cs == hipStreamCaptureStatusNone) {
    STEAM_HIP(hipEventRecord(b->uploaded, s));
    b->pending_upload = true;
  }
  ProfScope ps(STEAM_KF_NOISE, s);
  return launch_pass(b, BP_NOISE_EXT, s, 0);
}

namespace {

int advance_passes(steam_batch* b, hipStream_t s) {
  int rc;
  {
    ProfScope ps(STEAM_KF_NOISE, s);
    if ((rc = launch_pass(b, BP_NOISE, s, 0))) return rc;
  }
  {
    ProfScope ps(STEAM_KF_ZFWD, s);
    if ((rc = launch_pass(b, BP_Z, s, 0))) return rc;
  }
  {
    ProfScope ps(STEAM_KF_YPASS, s);
    if ((rc = launch_pass(b, BP_Y, s, 0))) return rc;
  }
  {
    ProfScope ps(STEAM_KF_XPASS, s);
    if ((rc = launch_pass(b, BP_X, s, 0))) return rc;
  }
  {
    ProfScope ps(STEAM_KF_YPASS, s);
    if ((rc = launch_pass(b, BP_Y, s, 1))) return rc;
  }
  {
    ProfScope ps(STEAM_KF_ZINV, s);
    if ((rc = launch_pass(b, BP_ZINV, s, 0))) return rc;
  }
  {
    ProfScope ps(STEAM_KF_FINALIZE, s);
    if ((rc = launch_pass(b, BP_FIN, s, 0))) return rc;
  }
  {
    ProfScope ps(STEAM_KF_POST, s);
    if ((rc = launch_pass(b, BP_POST, s, 0))) return rc;
  }
  return STEAM_OK;
}

}  // namespace

extern "C" int steam_batch_advance(steam_batch* b, steam_stream_t stream) {
  if (!b) return set_error(STEAM_EINVAL, "steam_batch_advance: NULL batch");
  hipStream_t s = (hipStream_t)stream;
  STEAM_HIP(hipMemcpyAsync(b->dev, b->host, (size_t)b->nslots * sizeof(PassDesc),
                           hipMemcpyHostToDevice, s));
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  STEAM_HIP(hipStreamIsCapturing(s, &cs));
  if (cs == hipStreamCaptureStatusNone) {
    STEAM_HIP(hipEventRecord(b->uploaded, s));
    b->pending_upload = true;
  }
  return advance_passes(b, s);
}

extern "C" int steam_batch_advance_after_noise(steam_batch* b, steam_stream_t stream) {
  if (!b) return set_error(STEAM_EINVAL, "steam_batch_advance_after_noise: NULL batch");
  if (b->groups[BP_NOISE_EXT].empty() || !b->groups[BP_NOISE].empty())
    return set_error(STEAM_EINVAL,
                     "steam_batch_advance_after_noise: every instance must read a noise "
                     "buffer (the descriptors come from the batch's steam_batch_noise)");
  return advance_passes(b, (hipStream_t)stream);
}
