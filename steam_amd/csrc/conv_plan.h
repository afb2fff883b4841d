// Host-side plan of the convolution / flux-advance operator and the
// launch sequence of its passes (shared by the single and batched entries).
#pragma once

#include "conv_kernels.h"

struct steam_conv_plan {
  steam::ConvGeom g;  // field operator
  steam::ConvGeom gk; // kernel-spectrum build (one block)
  int precision;
  float2* tw_dev;    // twiddle tables
  int64_t nx, ny, nz, kx, ky, kz;
  int64_t s_bytes, pz_off, po_off, led_off, nf_off, ws_bytes, spec_bytes;
  int ktile;         // columns of the kernel spectrum's x-pass tiles (0: [nzc][nx][ny])
  int ksym_shape;    // the kernel's shape admits the real, x-even spectrum format
  int64_t kstd_off;  // spectrum build area in [nzc][nx][ny] order
  int wg_noise;
  size_t lds_z, lds_zinv, lds_y, lds_x;
  size_t lds_zk, lds_yk, lds_xk;  // the kernel-spectrum build's (geometry gk)
  int wg_post;
};

namespace steam {

inline int64_t align256(int64_t v) { return (v + 255) & ~(int64_t)255; }

// Workgroup cap of a noise field generated ahead, beside other work:
// STEAM_NOISE_WGS from the environment when set (read once), else the
// compiled default.
int noise_wgs_cap();

NoiseDev make_noise_dev(const steam_noise_params* np);
int check_noise_params(const steam_noise_params* np);
DevLedger ledger_to_dev(const steam_flux_ledger* h);
void set_out_scale(PassDesc& d, const DevLedger& L);
void ledger_to_host(const DevLedger& L, steam_flux_ledger* h);

}  // namespace steam

namespace steam {

inline PassDesc base_desc(const steam_conv_plan* p, void* ws) {
  PassDesc d;
  memset(&d, 0, sizeof(d));
  d.g = p->g;
  char* w = (char*)ws;
  d.S = (float2*)w;
  d.part_z = (double*)(w + p->pz_off);
  d.part_o = (double*)(w + p->po_off);
  d.led = (DevLedger*)(w + p->led_off);
  d.nfield = (float*)(w + p->nf_off);
  d.kahan = p->precision == STEAM_PREC_KAHAN;
  d.win[0] = 0;
  d.win[1] = (int)p->nx;
  d.win[2] = 0;
  d.win[3] = (int)p->ny;
  d.sigma_in = 1.0f;
  return d;
}

#define STEAM_RS_DISPATCH(rs, ...)          \
  switch (rs) {                             \
    case 0: { constexpr int RS_ = 0; __VA_ARGS__; } break; \
    case 1: { constexpr int RS_ = 1; __VA_ARGS__; } break; \
    default: { constexpr int RS_ = 2; __VA_ARGS__; } break; \
  }

// the x/y passes add the Bluestein set
#define STEAM_RS_DISPATCH_XY(rs, ...)       \
  switch (rs) {                             \
    case 0: { constexpr int RS_ = 0; __VA_ARGS__; } break; \
    case 1: { constexpr int RS_ = 1; __VA_ARGS__; } break; \
    case 3: { constexpr int RS_ = 3; __VA_ARGS__; } break; \
    default: { constexpr int RS_ = 2; __VA_ARGS__; } break; \
  }

#define STEAM_LOGN_DISPATCH(logn, ...)                     \
  switch (logn) {                                          \
    case 6: { constexpr int LN_ = 6; __VA_ARGS__; } break;  \
    case 7: { constexpr int LN_ = 7; __VA_ARGS__; } break;  \
    case 8: { constexpr int LN_ = 8; __VA_ARGS__; } break;  \
    case 9: { constexpr int LN_ = 9; __VA_ARGS__; } break;  \
    case 10: { constexpr int LN_ = 10; __VA_ARGS__; } break; \
    default: { constexpr int LN_ = 11; __VA_ARGS__; } break; \
  }

template <class SrcT>
inline void launch_ypass(const ConvGeom& g, size_t lds_y, hipStream_t s,
                         const SrcT& src, int inv) {
  if (g.reg_y) {
    STEAM_LOGN_DISPATCH(g.reg_y, hipLaunchKernelGGL((ypass_reg_kernel<SrcT, LN_>),
                                                    dim3(g.wg_y), dim3(256), 0, s,
                                                    src, inv));
  } else if (g.mr_y) {
    with_xyplan(g.mr_y, [&](auto xp) {
      hipLaunchKernelGGL((ypass_mr_kernel<SrcT, decltype(xp)>), dim3(g.wg_y),
                         dim3(xy_nth<decltype(xp)>()), 0, s, src, inv);
    });
  } else {
    const int rsy = radix_set(g.ay);
    STEAM_RS_DISPATCH_XY(rsy, hipLaunchKernelGGL((ypass_kernel<SrcT, RS_>), dim3(g.wg_y),
                                                 dim3(generic_xy_nth(RS_)), lds_y, s, src, inv));
  }
}

template <class SrcT, int XMODE>
inline void launch_xpass(const ConvGeom& g, size_t lds_x, hipStream_t s,
                         const SrcT& src) {
  if (g.reg_x) {
    const int threads = (g.nx / 8 / xreg_lanes(g.reg_x)) * g.Tr;
    STEAM_LOGN_DISPATCH(g.reg_x, hipLaunchKernelGGL(
        (xpass_reg_kernel<SrcT, XMODE, LN_, xreg_lanes(LN_)>), dim3(g.wg_x), dim3(threads),
        0, s, src));
  } else if (g.mr_x) {
    with_xyplan(g.mr_x, [&](auto xp) {
      hipLaunchKernelGGL((xpass_mr_kernel<SrcT, XMODE, decltype(xp)>), dim3(g.wg_x),
                         dim3(xy_nth<decltype(xp)>()), 0, s, src);
    });
  } else {
    const int rsx = radix_set(g.ax);
    STEAM_RS_DISPATCH_XY(rsx, hipLaunchKernelGGL((xpass_kernel<SrcT, XMODE, RS_>),
                                                 dim3(g.wg_x), dim3(generic_xy_nth(RS_)), lds_x,
                                                 s, src));
  }
}

template <int MODE>
inline int run_operator(const steam_conv_plan* p, PassDesc& d, hipStream_t s) {
  const ConvGeom& g = d.g;
  const int rsz = radix_set(g.az);
  // the tile kernels move 16-byte vectors: every field pointer must allow it
  auto a16 = [](const void* q) { return ((uintptr_t)q & 15) == 0; };
  auto a8 = [](const void* q) { return ((uintptr_t)q & 7) == 0; };
  // the even-level register path moves 8-byte pairs; the odd one 4-byte values
  const bool odd = g.nz & 1;
  const bool zreg = g.zreg && (odd || (a8(d.in) && a8(d.out) && a8(d.nfield) && a8(d.sk)));
  const bool tile = !zreg && g.tile_z && a16(d.in) && a16(d.out) && a16(d.nfield) &&
                    a16(d.sk);
  d.npart = zreg ? g.wg_zr : g.wg_z;
  // the served flux options take the register z passes compiled with them
  const bool xf = MODE == MODE_FLUX && (d.unit_in || d.out_scale);
  SingleSrc src{d};
  {
    ProfScope ps(STEAM_KF_ZFWD, s);
    if (zreg) {
      with_zplan(g.n2, [&](auto zp) {
        using ZP = decltype(zp);
        if (xf && odd)
          hipLaunchKernelGGL((zfwd_reg_kernel<SingleSrc, MODE, ZP, true, true>),
                             dim3(g.wg_zr), dim3(ZR_NTH), 0, s, src);
        else if (xf)
          hipLaunchKernelGGL((zfwd_reg_kernel<SingleSrc, MODE, ZP, false, true>),
                             dim3(g.wg_zr), dim3(ZR_NTH), 0, s, src);
        else if (odd)
          hipLaunchKernelGGL((zfwd_reg_kernel<SingleSrc, MODE, ZP, true>),
                             dim3(g.wg_zr), dim3(ZR_NTH), 0, s, src);
        else
          hipLaunchKernelGGL((zfwd_reg_kernel<SingleSrc, MODE, ZP, false>),
                             dim3(g.wg_zr), dim3(ZR_NTH), 0, s, src);
      });
    } else if (tile) {
      STEAM_RS_DISPATCH(rsz, hipLaunchKernelGGL((zfwd_tile_kernel<SingleSrc, MODE, RS_>),
                                                dim3(g.wg_z), dim3(Z_NTH), p->lds_z, s, src));
    } else {
      STEAM_RS_DISPATCH(rsz, hipLaunchKernelGGL((zfwd_kernel<SingleSrc, MODE, RS_>),
                                                dim3(g.wg_z), dim3(Z_NTH), p->lds_z, s, src));
    }
  }
  STEAM_LAUNCH_CHECK("zfwd_kernel");
  {
    ProfScope ps(STEAM_KF_YPASS, s);
    launch_ypass(g, p->lds_y, s, src, 0);
  }
  STEAM_LAUNCH_CHECK("ypass_kernel");
  {
    ProfScope ps(STEAM_KF_XPASS, s);
    launch_xpass<SingleSrc, 0>(g, p->lds_x, s, src);
  }
  STEAM_LAUNCH_CHECK("xpass_kernel");
  {
    ProfScope ps(STEAM_KF_YPASS, s);
    launch_ypass(g, p->lds_y, s, src, 1);
  }
  STEAM_LAUNCH_CHECK("ypass_kernel(inv)");
  {
    ProfScope ps(STEAM_KF_ZINV, s);
    if (zreg) {
      with_zplan(g.n2, [&](auto zp) {
        using ZP = decltype(zp);
        if (xf && odd)
          hipLaunchKernelGGL((zinv_reg_kernel<SingleSrc, MODE, ZP, true, true>),
                             dim3(g.wg_zr), dim3(ZR_NTH), 0, s, src);
        else if (xf)
          hipLaunchKernelGGL((zinv_reg_kernel<SingleSrc, MODE, ZP, false, true>),
                             dim3(g.wg_zr), dim3(ZR_NTH), 0, s, src);
        else if (odd)
          hipLaunchKernelGGL((zinv_reg_kernel<SingleSrc, MODE, ZP, true>),
                             dim3(g.wg_zr), dim3(ZR_NTH), 0, s, src);
        else
          hipLaunchKernelGGL((zinv_reg_kernel<SingleSrc, MODE, ZP, false>),
                             dim3(g.wg_zr), dim3(ZR_NTH), 0, s, src);
      });
    } else if (tile) {
      STEAM_RS_DISPATCH(rsz, hipLaunchKernelGGL((zinv_tile_kernel<SingleSrc, MODE, RS_>),
                                                dim3(g.wg_z), dim3(Z_NTH), p->lds_zinv, s, src));
    } else {
      STEAM_RS_DISPATCH(rsz, hipLaunchKernelGGL((zinv_kernel<SingleSrc, MODE, RS_>),
                                                dim3(g.wg_z), dim3(Z_NTH), p->lds_zinv, s, src));
    }
  }
  STEAM_LAUNCH_CHECK("zinv_kernel");
  return STEAM_OK;
}

}  // namespace steam
