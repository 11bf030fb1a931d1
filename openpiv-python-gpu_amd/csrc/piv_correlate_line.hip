// Window correlation for FFT sizes F = 16 and 32 with one lane per line.
//
// Same semantics as corr_kernel (CorrelationGPU, gpu_process.py:1021-1310:
// window extraction :1420-1449, mean normalisation :1201-1205, zero padding
// :1222-1235, irfft2(conj(rfft2(a)) * rfft2(b)) :1237-1246, fftshift :1218,
// first-occurrence argmax :1248-1254, 3-point sub-pixel fits :1256-1294,
// u = j_sp - F/2, v = i_sp - F/2 :1296-1310), different mapping: every
// 1-D transform of length F runs entirely in the registers of one lane, so
// no transform needs a cross-lane butterfly; the row <-> column transposes go
// through LDS.  A wave owns 64 / (F/2) windows (F/2 lanes each) and never
// waits for other waves: all synchronisation is within the wave.
//
//   stage  lane j of a window holds                      LDS
//   R      row j of z = (a - mean) + i (b - mean)        (full windows: rows
//          -> A and B row spectra (DC, Nyquist in slot 0) loaded by their lane;
//                                                        else staging) -> AB
//   C      columns kx = j of A and B -> conj(FA) FB      AB -> registers
//          (slot-0 spectra through LDS for the DC/Nyquist product)
//   IC     inverse column kx of the product              -> R
//   IR     rows 2j, 2j+1 as one complex inverse          R -> plane
//   peak   first-occurrence argmax, sub-pixel fit        plane
//
// Used for launches with search window <= F/2 that need neither the S2N
// epilogue nor an interrogation window inside the search window.
#include <stdlib.h>

#include "piv_buffer.h"
#include "piv_fft.h"
#include "piv_kernels.h"

namespace piv {

namespace {

template <typename T>
__device__ __forceinline__ float load_px(const void* base, size_t idx) {
  return (float)(reinterpret_cast<const T*>(base)[idx]);
}

// Orders this wave's LDS accesses (a wave's LDS instructions execute in
// order; the fence keeps the compiler from moving accesses across).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <int F>
struct LineLayout {
  static constexpr int H = F / 2;
  static constexpr int G = 64 / H;  // windows per wave
  static constexpr int SF = F + 1;  // AB row stride (complex)
  static constexpr int SR = H + 1;  // R row stride (complex)
  static constexpr int SP = F + 1;  // plane row stride (float)
  // occupancy the register allocation is held to (LDS allows about this many)
  static constexpr int WAVES = F == 16 ? 4 : 2;
};

struct Peak {
  float v;
  int i;
};
__device__ __forceinline__ Peak better(Peak a, Peak b) {
  // larger value wins; equal values -> smaller flat index (first occurrence)
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

}  // namespace

size_t correlate_line_lds_floats(int fsize, int ss) {
  const size_t staging = 2ull * ss * (ss + 1);
  const size_t ab = 2ull * ss * (fsize + 1);
  const size_t r = 2ull * fsize * (fsize / 2 + 1);
  const size_t plane = (size_t)fsize * (fsize + 1);
  size_t m = staging;
  if (ab > m) m = ab;
  if (r > m) m = r;
  if (plane > m) m = plane;
  // a wave's windows sit RF floats apart, RF odd: complex values move as
  // dword pairs (ds_read2/ds_write2_b32, 4-byte aligned), so neighbouring
  // windows fall on banks of opposite parity and the even lane strides of the
  // transposes do not collide (measured at F = 32: RF = 57 mod 64 about 5 %
  // less time than any multiple of 4; odd residues are within 1 % of it)
  return m + (57 + 64 - m % 64) % 64;
}

// BATCHED: one launch over the windows of many fields (pairs); block b
// belongs to the last field whose block_prefix entry is <= b, so the field's
// arguments stay uniform over the block.  BW: waves per workgroup (a
// launch-shape choice: every window is computed by one wave with the same
// code whatever BW is; single-pair launches use 1, so the LDS granularity is
// one wave's windows).
template <int F, typename InT, bool BATCHED, int BW>
__global__ __launch_bounds__(64 * BW, LineLayout<F>::WAVES) void corr_line_kernel(CorrArgs a_in, CorrBatch batch,
                                                                            int n_windows_in, int region_floats) {
  using L = LineLayout<F>;
  constexpr int H = L::H, G = L::G, SF = L::SF, SR = L::SR, SP = L::SP;
  extern __shared__ float4 smem4[];
  float* smem = reinterpret_cast<float*>(smem4);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane / H, j = lane % H;

  // XCD-aware block order (consecutive hardware blocks go round-robin over 8
  // XCDs; each XCD gets a contiguous band of windows)
  int bid = blockIdx.x;
  if (!BATCHED || batch.xcd) {
    const int nb = gridDim.x;
    const int per = nb / 8;
    if (per > 0 && bid < per * 8) bid = (bid & 7) * per + (bid >> 3);
  }
  CorrArgs a;
  int n_windows;
  if constexpr (BATCHED) {
    int lo = 0, hi = batch.n_fields - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (batch.block_prefix[mid] <= bid) lo = mid;
      else hi = mid - 1;
    }
    a = batch.fields[lo];
    bid -= batch.block_prefix[lo];
    n_windows = a.m * a.n;
  } else {
    a = a_in;
    n_windows = n_windows_in;
  }
  const int win = (bid * BW + wave) * G + g;
  const bool valid = win < n_windows;
  const int W = a.ss;
  // the predictor of the window, loaded now so the epilogue does not wait
  float pred_u = 0.f, pred_v = 0.f;
  if (a.up && j == 0 && valid) {
    pred_u = a.up[win];
    pred_v = a.vp[win];
  }
  const int pad = (F - W) / 2;
  const int SW = W + 1;

  float* region = smem + (size_t)(wave * G + g) * region_floats;
  float* stA = region;
  float* stB = region + W * SW;
  c32* AB = reinterpret_cast<c32*>(region);
  c32* R = reinterpret_cast<c32*>(region);
  float* plane = region;

  const int wrow = valid ? (win / a.n) : 0;
  const int wcol = valid ? (win % a.n) : 0;
  const int row0 = wrow * a.spacing + a.y_off;
  const int col0 = wcol * a.spacing + a.x_off;

  // ---- staging: the window's lanes load W x W pixels ----------------------
  // Full windows (the usual case): lane j loads row j itself, which is what
  // the row stage needs, so the pixels never pass through LDS.  Otherwise the
  // window goes through an LDS staging area, lane j reading row j back.
  float sa = 0.f, sb = 0.f;
  const bool direct = W == H && buffer_fits((uint64_t)a.ht * a.wd, sizeof(InT));
  float rowa[H], rowb[H];
  if (direct) {
    const uint32_t npx = (uint32_t)a.ht * (uint32_t)a.wd;
    const rsrc_t ra = make_rsrc(a.frame_a, npx * (uint32_t)sizeof(InT));
    const rsrc_t rb = make_rsrc(a.frame_b, npx * (uint32_t)sizeof(InT));
    const uint32_t base = (uint32_t)(row0 + j) * (uint32_t)a.wd + (uint32_t)col0;
#pragma unroll
    for (int c = 0; c < H; ++c) {
      rowa[c] = buf_px<InT>(ra, base + (uint32_t)c);
      rowb[c] = buf_px<InT>(rb, base + (uint32_t)c);
    }
#pragma unroll
    for (int c = 0; c < H; ++c) {
      rowa[c] = valid ? rowa[c] : 0.f;
      rowb[c] = valid ? rowb[c] : 0.f;
      sa += rowa[c];
      sb += rowb[c];
    }
  } else {
    int r = j / W, c = j % W;
    for (int e = j; e < W * W; e += H) {
      float va = 0.f, vb = 0.f;
      if (valid) {
        const size_t gi = (size_t)(row0 + r) * a.wd + (col0 + c);
        va = load_px<InT>(a.frame_a, gi);
        vb = load_px<InT>(a.frame_b, gi);
      }
      stA[r * SW + c] = va;
      stB[r * SW + c] = vb;
      sa += va;
      sb += vb;
      c += H;
      while (c >= W) {
        c -= W;
        r += 1;
      }
    }
    wave_lds_sync();
    const int rr = j < W ? j : 0;
#pragma unroll
    for (int c2 = 0; c2 < H; ++c2) {
      const bool in = c2 < W;
      const int ci = in ? c2 : 0;
      rowa[c2] = in ? stA[rr * SW + ci] : 0.f;
      rowb[c2] = in ? stB[rr * SW + ci] : 0.f;
    }
  }
#pragma unroll
  for (int off = H / 2; off > 0; off >>= 1) {
    sa += __shfl_xor(sa, off, 64);
    sb += __shfl_xor(sb, off, 64);
  }
  float mean_a = 0.f, mean_b = 0.f;
  if (a.normalize || pad != 0) {  // gpu_process.py:1072-1081, 1201-1205
    mean_a = sa / (float)(W * W);
    mean_b = sb / (float)(W * W);
  }
  wave_lds_sync();

  // The windows sit at index 0 of the F-point transforms instead of at the
  // padding offset (F - W)/2: a circular shift common to both frames cancels
  // in conj(FA) FB, so the plane is the same, and inputs F/2..F-1 are
  // compile-time zeros (W <= F/2).

  // ---- stage R: row j of z, all F points in this lane ---------------------
  const bool row_on = j < W;
  {
    c32 x[F];
    const int rr = row_on ? j : 0;
    static_for<0, H>([&](auto CC) {
      constexpr int c = decltype(CC)::value;
      const bool in = c < W;
      x[c] = in ? c32{rowa[c] - mean_a, rowb[c] - mean_b} : c32{0.f, 0.f};
    });
    static_for<H, F>([&](auto CC) { x[decltype(CC)::value] = c32{0.f, 0.f}; });
    wave_lds_sync();  // staging (aliased by AB) fully read
    fft_reg<F, -1, true>(x);
    // Z[k] sits in x[bitrev(k)].  2A[k] = Z[k] + conj(Z[F-k]),
    // 2B[k] = -i (Z[k] - conj(Z[F-k])); A and B are real at k = 0 and k = H,
    // packed as slot 0 = (DC, Nyquist).  The factors 2 are folded into the
    // plane scale.
    if (row_on) {
      c32* row = AB + (size_t)rr * SF;
      const c32 z0 = x[0], zh = x[bitrev(H, F)];
      row[0] = c32{2.f * z0.x, 2.f * zh.x};
      row[H] = c32{2.f * z0.y, 2.f * zh.y};
      static_for<1, H>([&](auto KK) {
        constexpr int k = decltype(KK)::value;
        const c32 zk = x[bitrev(k, F)], zm = x[bitrev(F - k, F)];
        row[k] = c32{zk.x + zm.x, zk.y - zm.y};
        const c32 d{zk.x - zm.x, zk.y + zm.y};
        row[H + k] = c32{d.y, -d.x};
      });
    }
  }
  wave_lds_sync();

  // ---- stage C: columns kx = j of A and B, product -------------------------
  // FA stays in registers while FB is transformed (occupancy is set by LDS,
  // not registers); lane 0 of the window copies its slot-0 spectra (FA to
  // column 0, FB to column H of the R layout) for the DC/Nyquist product.
  const int kx = j;
  c32 p[F];  // product P[ky][kx] in natural order of ky
  {
    c32 bin[H];
    static_for<0, H>([&](auto MM) {
      constexpr int r = decltype(MM)::value;
      const bool in = r < W;
      const int ri = in ? r : 0;
      const c32 va = AB[(size_t)ri * SF + kx], vb = AB[(size_t)ri * SF + H + kx];
      p[r] = in ? va : c32{0.f, 0.f};
      bin[r] = in ? vb : c32{0.f, 0.f};
    });
    static_for<H, F>([&](auto MM) { p[decltype(MM)::value] = c32{0.f, 0.f}; });
    wave_lds_sync();  // AB (aliased by R) fully read
    fft_reg<F, -1, true>(p);  // FA of column kx, frequency ky in p[bitrev(ky)]
    c32 fb[F];
    static_for<0, H>([&](auto MM) { fb[decltype(MM)::value] = bin[decltype(MM)::value]; });
    static_for<H, F>([&](auto MM) { fb[decltype(MM)::value] = c32{0.f, 0.f}; });
    fft_reg<F, -1, true>(fb);
    if (kx == 0) {
      static_for<0, F>([&](auto KY) {
        constexpr int ky = decltype(KY)::value;
        R[(size_t)ky * SR] = p[bitrev(ky, F)];
        R[(size_t)ky * SR + H] = fb[bitrev(ky, F)];
      });
    }
    // conj(FA) FB in place, then back to natural order of ky
    static_for<0, F>([&](auto KY) {
      constexpr int ky = decltype(KY)::value;
      p[ky] = cconjmul(p[ky], fb[ky]);
    });
    static_for<0, F>([&](auto KY) {
      constexpr int ky = decltype(KY)::value;
      constexpr int b = bitrev(ky, F);
      if constexpr (ky < b) {
        const c32 tmp = p[ky];
        p[ky] = p[b];
        p[b] = tmp;
      }
    });
  }
  wave_lds_sync();
  {
    // Slot 0 holds FFT(col_DC + i col_Nyquist) of two real columns: unpack
    // with the ky <-> -ky mirror, form both products and repack as
    // P_DC + i P_Nyquist (both invert to real columns).  Lane j of the window
    // does ky = j and ky = j + F/2.
    c32 q2[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int ky = j + t * H, km = (F - ky) & (F - 1);
      const c32 qa = R[(size_t)ky * SR], qam = R[(size_t)km * SR];
      const c32 qb = R[(size_t)ky * SR + H], qbm = R[(size_t)km * SR + H];
      // f0 = (q + conj(qm)) / 2 ; fn = (q - conj(qm)) / (2i)
      const c32 a0{0.5f * (qa.x + qam.x), 0.5f * (qa.y - qam.y)};
      const c32 an{0.5f * (qa.y + qam.y), 0.5f * (qam.x - qa.x)};
      const c32 b0{0.5f * (qb.x + qbm.x), 0.5f * (qb.y - qbm.y)};
      const c32 bn{0.5f * (qb.y + qbm.y), 0.5f * (qbm.x - qb.x)};
      const c32 p0 = cconjmul(a0, b0), pn = cconjmul(an, bn);
      q2[t] = c32{p0.x - pn.y, p0.y + pn.x};
    }
    wave_lds_sync();  // every mirror read done before column H is overwritten
    R[(size_t)j * SR + H] = q2[0];
    R[(size_t)(j + H) * SR + H] = q2[1];
    wave_lds_sync();
    if (kx == 0) {
      static_for<0, F>([&](auto KY) {
        constexpr int ky = decltype(KY)::value;
        p[ky] = R[(size_t)ky * SR + H];
      });
    }
  }

  // ---- stage IC: inverse column kx ------------------------------------------
  // output row y goes to slot (y >> 1) + H (y & 1): the rows 2j, 2j+1 that
  // lane j reads next sit in slots j and j + H, an odd stride between lanes
  fft_reg<F, +1>(p);
  wave_lds_sync();  // lane 0's reads of column H done before R is written
  static_for<0, F>([&](auto YY) {
    constexpr int y = decltype(YY)::value;
    R[(size_t)((y >> 1) + H * (y & 1)) * SR + kx] = p[bitrev(y, F)];
  });
  wave_lds_sync();

  // ---- stage IR: rows 2j and 2j+1 as real and imaginary part ---------------
  const int y1 = 2 * j;  // rows y1 and y1 + 1 (slots j and j + H)
  c32 q[F];
  {
    c32 l1[H], l2[H];
    static_for<0, H>([&](auto SS) {
      constexpr int s = decltype(SS)::value;
      l1[s] = R[(size_t)j * SR + s];
      l2[s] = R[(size_t)(j + H) * SR + s];
    });
    // row spectra are Hermitian: S[F - k] = conj(S[k]); slot 0 = (DC, Nyquist)
    q[0] = c32{l1[0].x, l2[0].x};
    q[H] = c32{l1[0].y, l2[0].y};
    static_for<1, H>([&](auto KK) {
      constexpr int k = decltype(KK)::value;
      q[k] = c32{l1[k].x - l2[k].y, l1[k].y + l2[k].x};
      q[F - k] = c32{l1[k].x + l2[k].y, l2[k].x - l1[k].y};
    });
  }
  wave_lds_sync();  // R (aliased by the plane) fully read
  fft_reg<F, +1>(q);

  // ---- plane (fftshift-ed) and peak ------------------------------------------
  const float scale = 1.0f / (float)(4 * F * F);  // 4: the split's factors of 2
  Peak best;
  {
    const int i1 = y1 ^ H, i2 = i1 + 1;  // y1 even: rows i1 < i2
    float* gp = (a.planes && valid) ? a.planes + (size_t)win * F * F : nullptr;
    // maximum on the raw values (the scale is a power of two), four partial
    // maxima so the comparisons do not form one chain
    float pm[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    static_for<0, F>([&](auto XX) {
      constexpr int xx = decltype(XX)::value;
      constexpr int jj = xx ^ H;
      const c32 val = q[bitrev(xx, F)];
      const float c1 = val.x * scale, c2 = val.y * scale;
      plane[i1 * SP + jj] = c1;
      plane[i2 * SP + jj] = c2;
      pm[xx & 3] = fmaxf(pm[xx & 3], fmaxf(val.x, val.y));
      if (gp) {
        gp[i1 * F + jj] = c1;
        gp[i2 * F + jj] = c2;
      }
    });
    const float qmax = fmaxf(fmaxf(pm[0], pm[1]), fmaxf(pm[2], pm[3]));
    // first occurrence of the maximum: flat index i * F + jj, here as the code
    // row * F + jj (row 0 = i1, 1 = i2), found by four independent chains
    // (rows x column halves) that walk their columns downwards with
    // compile-time codes; the smallest code wins
    int fc[4] = {2 * F, 2 * F, 2 * F, 2 * F};
    static_for<0, H>([&](auto JJ) {
      constexpr int jl = H - 1 - decltype(JJ)::value;  // lower column half
      constexpr int jh = jl + H;                         // upper column half
      fc[0] = q[bitrev(jl ^ H, F)].x == qmax ? jl : fc[0];
      fc[1] = q[bitrev(jh ^ H, F)].x == qmax ? jh : fc[1];
      fc[2] = q[bitrev(jl ^ H, F)].y == qmax ? F + jl : fc[2];
      fc[3] = q[bitrev(jh ^ H, F)].y == qmax ? F + jh : fc[3];
    });
    const int code = min(min(fc[0], fc[1]), min(fc[2], fc[3]));
    const int first = code < 2 * F ? i1 * F + code : 0x7fffffff;  // i2 = i1 + 1
    best = Peak{qmax * scale, first};
  }
#pragma unroll
  for (int off = H / 2; off > 0; off >>= 1) {
    const Peak o{__shfl_xor(best.v, off, 64), __shfl_xor(best.i, off, 64)};
    best = better(best, o);
  }
  wave_lds_sync();

  if (j == 0 && valid) {
    const int ip = best.i / F;
    const int jp = best.i % F;
    float isp = (float)ip, jsp = (float)jp;
    if (ip > 0 && ip < F - 1 && jp > 0 && jp < F - 1) {
      const float c = plane[ip * SP + jp];
      const float cd = plane[(ip - 1) * SP + jp];
      const float cu = plane[(ip + 1) * SP + jp];
      const float cl = plane[ip * SP + jp - 1];
      const float cr = plane[ip * SP + jp + 1];
      const float eps = 1e-7f;
      if (a.subpixel == kGaussian) {
        if (c > 0.f && cd > 0.f && cu > 0.f && cl > 0.f && cr > 0.f) {
          const float lc = logf(c), lcd = logf(cd), lcu = logf(cu), lcl = logf(cl), lcr = logf(cr);
          isp = (float)ip + 0.5f * (lcd - lcu) / (lcd - 2.0f * lc + lcu + eps);
          jsp = (float)jp + 0.5f * (lcl - lcr) / (lcl - 2.0f * lc + lcr + eps);
        }
      } else if (a.subpixel == kParabolic) {
        isp = (float)ip + 0.5f * (cd - cu) / (cd - 2.0f * c + cu + eps);
        jsp = (float)jp + 0.5f * (cl - cr) / (cl - 2.0f * c + cr + eps);
      } else {
        isp = (float)ip + (cu - cd) / (cd + c + cu + eps);
        jsp = (float)jp + (cr - cl) / (cl + c + cr + eps);
      }
    }
    float uu = jsp - (float)(F / 2);
    float vv = isp - (float)(F / 2);
    if (a.up) {
      uu += pred_u;
      vv += pred_v;
    }
    a.u[win] = uu;
    a.v[win] = vv;
    if (a.i_peak) {
      a.i_peak[win] = ip;
      a.j_peak[win] = jp;
    }
  }
}

namespace {

template <int F, typename InT, bool BATCHED, int BW>
hipError_t line_lds_attr(size_t lds) {
  static thread_local size_t attr_bytes = 0;
  if (lds > 65536 && lds > attr_bytes) {
    hipError_t e = hipFuncSetAttribute((const void*)corr_line_kernel<F, InT, BATCHED, BW>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
      set_launch_detail("hipFuncSetAttribute(MaxDynamicSharedMemorySize)", e);
      return e;
    }
    attr_bytes = lds;
  }
  return hipSuccess;
}

template <int F, typename InT>
hipError_t launch_line_t(const CorrArgs& a, hipStream_t stream) {
  using L = LineLayout<F>;
  const int n_windows = a.m * a.n;
  const int region = (int)correlate_line_lds_floats(F, a.ss);
  const size_t lds = (size_t)region * sizeof(float) * L::G;
  const int per_block = L::G;
  const int blocks = (n_windows + per_block - 1) / per_block;
  (void)hipGetLastError();
  hipError_t e = line_lds_attr<F, InT, false, 1>(lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((corr_line_kernel<F, InT, false, 1>), dim3(blocks), dim3(64), lds, stream, a, CorrBatch{},
                     n_windows, region);
  e = hipGetLastError();
  if (e != hipSuccess) set_launch_detail("corr_line_kernel launch", e);
  return e;
}

template <int F, typename InT, int BW>
hipError_t launch_line_batch_t(const CorrBatch& b, int total_blocks, int max_ss, hipStream_t stream) {
  using L = LineLayout<F>;
  const int region = (int)correlate_line_lds_floats(F, max_ss);
  const size_t lds = (size_t)region * sizeof(float) * L::G * BW;
  (void)hipGetLastError();
  hipError_t e = line_lds_attr<F, InT, true, BW>(lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((corr_line_kernel<F, InT, true, BW>), dim3(total_blocks), dim3(64 * BW), lds, stream, CorrArgs{},
                     b, 0, region);
  e = hipGetLastError();
  if (e != hipSuccess) set_launch_detail("corr_line_kernel (batched) launch", e);
  return e;
}

template <int F>
hipError_t launch_line_f(const CorrArgs& a, hipStream_t stream) {
  switch (a.in_dtype) {
    case kU8: return launch_line_t<F, uint8_t>(a, stream);
    case kU16: return launch_line_t<F, uint16_t>(a, stream);
    case kF32: return launch_line_t<F, float>(a, stream);
  }
  return hipErrorInvalidValue;
}

template <int F, int BW>
hipError_t launch_line_batch_fw(const CorrBatch& b, int in_dtype, int total_blocks, int max_ss, hipStream_t stream) {
  switch (in_dtype) {
    case kU8: return launch_line_batch_t<F, uint8_t, BW>(b, total_blocks, max_ss, stream);
    case kU16: return launch_line_batch_t<F, uint16_t, BW>(b, total_blocks, max_ss, stream);
    case kF32: return launch_line_batch_t<F, float, BW>(b, total_blocks, max_ss, stream);
  }
  return hipErrorInvalidValue;
}
template <int F>
hipError_t launch_line_batch_f(const CorrBatch& b, int in_dtype, int total_blocks, int max_ss, int bw,
                               hipStream_t stream) {
  switch (bw) {
    case 1: return launch_line_batch_fw<F, 1>(b, in_dtype, total_blocks, max_ss, stream);
    case 2: return launch_line_batch_fw<F, 2>(b, in_dtype, total_blocks, max_ss, stream);
    case 4: return launch_line_batch_fw<F, 4>(b, in_dtype, total_blocks, max_ss, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace

bool correlate_line_enabled() {
  static const bool on = [] {
    const char* e = getenv("OPENPIV_HIP_CORR_LINE");
    return !(e && e[0] == '0');
  }();
  return on;
}

int correlate_line_windows_per_block(int fsize, int bw) {
  return (fsize == 16 ? LineLayout<16>::G : LineLayout<32>::G) * bw;
}

hipError_t launch_correlate_line_batch(const CorrBatch& b, int fsize, int in_dtype, int total_blocks, int max_ss,
                                       int bw, hipStream_t stream) {
  if (total_blocks <= 0) return hipSuccess;
  switch (fsize) {
    case 16: return launch_line_batch_f<16>(b, in_dtype, total_blocks, max_ss, bw, stream);
    case 32: return launch_line_batch_f<32>(b, in_dtype, total_blocks, max_ss, bw, stream);
  }
  return hipErrorInvalidValue;
}

// F = 64 is left to corr_kernel: its 17 KB of LDS per window would hold this
// mapping to one wave per SIMD, which measured slower than corr_kernel.
bool correlate_line_applies(const CorrArgs& a) {
  return (a.fsize == 16 || a.fsize == 32) && 2 * a.ss <= a.fsize && a.s2n == nullptr && a.win_pad == 0 &&
         (a.in_dtype == kU8 || a.in_dtype == kU16 || a.in_dtype == kF32);
}

hipError_t launch_correlate_line(const CorrArgs& a, hipStream_t stream) {
  switch (a.fsize) {
    case 16: return launch_line_f<16>(a, stream);
    case 32: return launch_line_f<32>(a, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace piv
