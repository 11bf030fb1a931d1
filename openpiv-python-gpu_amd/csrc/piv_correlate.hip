// Fused window correlation for one PIV pass (gfx950, wave64).
//
// Semantics follow CorrelationGPU (gpu_process.py:1021-1310): windows are
// extracted on the pass grid (cuda_get_stack, :1420-1449), mean-normalised
// over ss x ss (:1201-1205, forced on whenever padding is nonzero, :1065),
// zero-padded to F x F with (F-ss)/2 on each side (:1222-1235), correlated
// as irfft2(conj(rfft2(a)) * rfft2(b)) (:1237-1246), fftshift-ed (:1218),
// reduced to a first-occurrence argmax (:1248-1254) and a 3-point sub-pixel
// fit (:1256-1294), and turned into u = j_sp - F/2, v = i_sp - F/2
// (:1296-1310).  No plane ever touches HBM.
//
// Mapping: 2F lanes per window, 256 threads per workgroup.  One 1-D FFT of
// length F lives in 2 lanes (F/2 points each, forward rows/columns) or in 4
// lanes (F/4 points each, inverse columns/rows); LDS only carries the two
// transposes and the final plane.
//
//   stage  what                                   lanes per FFT   LDS
//   R      rows of z = a + i b (ss nonzero rows)  2               staging -> AB
//   S      split z-row spectra into A and B rows  (in R's lanes)
//   C      columns of A and B (one quad per kx)   2               AB -> regs
//   P      P = conj(FA) FB, DC/Nyquist packed     (quad)
//   IC     inverse columns of P                   4               regs -> R
//   IR     inverse rows, two rows per transform   4               R -> plane
//   peak   argmax + sub-pixel                                     plane
#include "piv_fft.h"
#include "piv_kernels.h"

namespace piv {

namespace {

template <typename T>
__device__ __forceinline__ float load_px(const void* base, size_t idx) {
  return (float)(reinterpret_cast<const T*>(base)[idx]);
}

struct Best {
  float v;
  int i;
};
__device__ __forceinline__ Best better(Best a, Best b) {
  // larger value wins; equal values -> smaller flat index (first occurrence)
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

template <int F>
struct Layout {
  static constexpr int H = F / 2;
  static constexpr int Q = F / 4;
  static constexpr int LANES = 2 * F;
  static constexpr int WPB = 256 / LANES;  // windows per workgroup
  static constexpr int SF = F + 1;         // AB row stride (complex)
  static constexpr int SR = H + 1;         // R row stride (complex)
  static constexpr int SP = F + 1;         // plane row stride (float)
  // occupancy target (waves per SIMD) the register allocation is held to
  // (measured best per size; F=128 fits two 66.5 KB windows per CU)
  static constexpr int WAVES = F <= 32 ? 8 : (F == 64 ? 3 : 2);
};

}  // namespace

size_t correlate_lds_bytes(int fsize, int ss) {
  size_t staging = 2ull * ss * (ss + 1) * sizeof(float);
  size_t ab = (size_t)ss * (fsize + 1) * 8;
  size_t r = (size_t)fsize * (fsize / 2 + 1) * 8;
  size_t plane = (size_t)fsize * (fsize + 1) * sizeof(float);
  size_t m = staging;
  if (ab > m) m = ab;
  if (r > m) m = r;
  if (plane > m) m = plane;
  return (m + 15) & ~size_t(15);
}

// BATCHED: one launch over the windows of many fields (pairs); block b
// belongs to the last field whose block_prefix entry is <= b, so every block
// works on one field and the field's arguments stay wave-uniform.
// EXTRA: the S2N epilogue and the interrogation window inside the search
// window are compiled in (launches that need neither use the lean variant).
// HALF: search window <= F/2.  The window then sits at index 0 of the
// transforms instead of at the padding offset (F - W)/2 -- a circular shift
// common to both frames that cancels in conj(FA) FB -- so the upper half of
// every forward input is a compile-time zero.
template <int F, typename InT, bool BATCHED, bool EXTRA, bool HALF>
__global__ __launch_bounds__(256, Layout<F>::WAVES) void corr_kernel(CorrArgs a_in, CorrBatch batch,
                                                                     int n_windows_in, int region_floats) {
  using L = Layout<F>;
  constexpr int H = L::H, Q = L::Q, LANES = L::LANES, WPB = L::WPB;
  constexpr int SF = L::SF, SR = L::SR, SP = L::SP;

  extern __shared__ float4 smem4[];
  float* smem = reinterpret_cast<float*>(smem4);
  __shared__ float red_f[8][2];
  __shared__ Best red_b[8];
  __shared__ Best red_s[8];

  const int tid = threadIdx.x;
  const int wl = tid / LANES;  // window slot in this workgroup
  const int t = tid % LANES;   // lane within the window
  const int wave = tid >> 6;

  // XCD-aware block order: consecutive hardware blocks go round-robin over 8
  // XCDs; give each XCD a contiguous band of windows so overlapping windows
  // share an L2.
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
  const int win = bid * WPB + wl;
  const bool valid = win < n_windows;
  const int W = a.ss;
  const int pad = (F - W) / 2;
  const int SW = W + 1;
  const int wp = EXTRA ? a.win_pad : 0;  // (search - interrogation window) / 2

  float* region = smem + (size_t)wl * region_floats;
  float* stA = region;
  float* stB = region + W * SW;
  c32* AB = reinterpret_cast<c32*>(region);
  c32* R = reinterpret_cast<c32*>(region);
  float* plane = region;

  const int wrow = valid ? (win / a.n) : 0;
  const int wcol = valid ? (win % a.n) : 0;
  const int row0 = wrow * a.spacing + a.y_off;
  const int col0 = wcol * a.spacing + a.x_off;

  // ---- staging: coalesced window load + partial sums --------------------
  float sa = 0.f, sb = 0.f;
  {
    const int total = W * W;
    int r = t / W, c = t % W;
    const int dr = LANES / W, dc = LANES % W;
    for (int e = t; e < total; e += LANES) {
      float va = 0.f, vb = 0.f;
      if (valid) {
        size_t gi = (size_t)(row0 + r) * a.wd + (col0 + c);
        // frame A keeps only the interrogation window inside the search
        // window (first pass with shrink_ratio, gpu_process.py:1066-1070)
        if (r >= wp && r < W - wp && c >= wp && c < W - wp) va = load_px<InT>(a.frame_a, gi);
        vb = load_px<InT>(a.frame_b, gi);
      }
      stA[r * SW + c] = va;
      stB[r * SW + c] = vb;
      sa += va;
      sb += vb;
      r += dr;
      c += dc;
      if (c >= W) {
        c -= W;
        r += 1;
      }
    }
  }
  // window-group sums
  constexpr int WIDTH = LANES < 64 ? LANES : 64;
#pragma unroll
  for (int off = WIDTH / 2; off > 0; off >>= 1) {
    sa += __shfl_xor(sa, off, 64);
    sb += __shfl_xor(sb, off, 64);
  }
  if constexpr (LANES > 64) {
    if ((tid & 63) == 0) {
      red_f[wave][0] = sa;
      red_f[wave][1] = sb;
    }
  }
  __syncthreads();
  if constexpr (LANES > 64) {
    constexpr int WPW = LANES / 64;  // waves per window
    const int w0 = (tid / LANES) * WPW;
    sa = 0.f;
    sb = 0.f;
    for (int k = 0; k < WPW; ++k) {
      sa += red_f[w0 + k][0];
      sb += red_f[w0 + k][1];
    }
  }
  float mean_a = 0.f, mean_b = 0.f;
  if (a.normalize || pad != 0) {
    // means over the interrogation window (A) and the search window (B)
    // (gpu_process.py:1072-1081, 1201-1205)
    const int wi = W - 2 * wp;
    mean_a = sa / (float)(wi * wi);
    mean_b = sb / (float)(W * W);
  }

  // ---- stage R: row FFTs of z = (a - mean_a) + i (b - mean_b) -----------
  c32 x[H];
  {
    const int q = t >> 1;
    // Opaque copy of the lane bit: the per-register bounds tests below and in
    // the column stage would otherwise be shared between the stages and held
    // in (spilled) lane masks across the barriers.
    int h = t & 1;
    asm volatile("" : "+v"(h));
    const bool active = q < W;
    const int rr = active ? q : 0;
    static_for<0, H>([&](auto M) {
      constexpr int m = decltype(M)::value;
      const int c = 2 * m + h - (HALF ? 0 : pad);
      c32 zz{0.f, 0.f};
      if ((!HALF || m < H / 2) && active && c >= 0 && c < W) {
        const bool in_a = rr >= wp && rr < W - wp && c >= wp && c < W - wp;
        zz.x = in_a ? stA[rr * SW + c] - mean_a : 0.f;
        zz.y = stB[rr * SW + c] - mean_b;
      }
      x[m] = zz;
    });
  }
  __syncthreads();  // staging (aliased by AB) fully consumed
  {
    const int q = t >> 1;
    const bool up1 = t & 1;
    fft_reg<H, -1>(x);
    // the last DIT stage across the lane pair (dit2_combine) is fused with
    // the split, so each step holds only the two registers it splits
    auto combine = [&](auto KK) {
      constexpr int K = decltype(KK)::value;
      constexpr int r = bitrev(K, H);
      const c32 own = x[r];
      const c32 tw = twiddle<K, F, -1>(own);
      const c32 v = up1 ? tw : own;
      const c32 recv = swap1(v);
      const c32 sv = up1 ? c32{-v.x, -v.y} : v;
      x[r] = c32{recv.x + sv.x, recv.y + sv.y};
    };
    // split: lane 0 -> A slots, lane 1 -> B slots; slot 0 packs (DC, Nyquist)
    const bool active = q < W;
    c32* row = AB + (size_t)(active ? q : 0) * SF + (up1 ? H : 0);
    {
      combine(std::integral_constant<int, 0>{});
      c32 o = x[0];
      c32 w = swap1(o);
      c32 z0 = up1 ? w : o;  // Zr[0]
      c32 zn = up1 ? o : w;  // Zr[F/2]
      c32 sv = up1 ? c32{2.f * z0.y, 2.f * zn.y} : c32{2.f * z0.x, 2.f * zn.x};
      if (active) row[0] = sv;
    }
    // slots kx and H-kx read the same two registers: one pair of swaps
    // serves both.  Lane 1 turns its inputs by -i so that both lanes use
    // the A formula: B(zk, zm) = A(-i zk, -i zm).  The 1/2 of the split is
    // folded into the final plane scale.
    static_for<1, H / 2 + 1>([&](auto K) {
      constexpr int kx = decltype(K)::value;
      constexpr int kx2 = H - kx;
      constexpr int r1 = bitrev(kx, H);
      constexpr int r2 = bitrev(kx2, H);
      combine(std::integral_constant<int, kx>{});
      if constexpr (kx2 != kx) combine(std::integral_constant<int, kx2>{});
      const c32 o1 = x[r1], w1 = swap1(o1);
      const c32 o2 = x[r2], w2 = swap1(o2);
      // slot kx : Zr[kx] = (o1 | w1), Zr[F-kx] = (w2 | o2)
      // slot kx2: Zr[kx2] = (o2 | w2), Zr[F-kx2] = (w1 | o1)
      const c32 zk1 = up1 ? c32{w1.y, -w1.x} : o1;
      const c32 zm1 = up1 ? c32{o2.y, -o2.x} : w2;
      const c32 s1{zk1.x + zm1.x, zk1.y - zm1.y};
      if (active) row[kx] = s1;
      if constexpr (kx2 != kx) {
        const c32 zk2 = up1 ? c32{w2.y, -w2.x} : o2;
        const c32 zm2 = up1 ? c32{o1.y, -o1.x} : w1;
        const c32 s2{zk2.x + zm2.x, zk2.y - zm2.y};
        if (active) row[kx2] = s2;
      }
    });
  }
  __syncthreads();

  // ---- stage C: column FFTs (quad = {A h0, A h1, B h0, B h1} per kx) ------
  const int kx = t >> 2;
  const int sel = (t >> 1) & 1;  // 0 -> A column, 1 -> B column
  int upper_bit = t & 1;
  asm volatile("" : "+v"(upper_bit));  // see the row stage
  const bool upper = upper_bit;
  {
    const int col = sel * H + kx;
    static_for<0, H>([&](auto M) {
      constexpr int m = decltype(M)::value;
      const int r = 2 * m + (upper ? 1 : 0) - (HALF ? 0 : pad);
      c32 zz{0.f, 0.f};
      if ((!HALF || m < H / 2) && r >= 0 && r < W) zz = AB[(size_t)r * SF + col];
      x[m] = zz;
    });
  }
  __syncthreads();  // AB fully consumed (R aliases it)
  fft_reg<H, -1>(x);
  dit2_combine<F, -1>(x, upper);

  // ---- stage P: spectrum product -------------------------------------------
  // Column slot 0 holds FFT(col0 + i colN) of the real DC and Nyquist columns
  // and needs the ky -> -ky mirror; its quad hands the two columns to the
  // window's first F threads through LDS (one ky each) instead of working
  // through all of them while the rest of the window waits.
  c32* xq = reinterpret_cast<c32*>(smem + (size_t)WPB * region_floats) + (size_t)wl * 3 * F;
  if (kx == 0) {
    static_for<0, H>([&](auto KK) {
      constexpr int k = decltype(KK)::value;
      xq[sel * F + k + (upper ? H : 0)] = x[bitrev(k, H)];
    });
  }
  {
    // A lanes: conj(own) * oth = conj(FA) FB; B lanes get its conjugate
    const float sgn = sel ? -1.f : 1.f;
    static_for<0, H>([&](auto RR) {
      constexpr int r = decltype(RR)::value;
      const c32 own = x[r];
      const c32 oth = swap2(own);
      const c32 c = cconjmul(own, oth);
      x[r] = c32{c.x, sgn * c.y};
    });
  }
  __syncthreads();
  if (t < F) {
    // unpack with the mirror: f0 = (q + conj(qm)) / 2, fn = (q - conj(qm)) / (2i);
    // both products, repacked as q = P0 + i PN (both columns invert to real data)
    const int ky = t, km = (F - t) & (F - 1);
    const c32 qa = xq[ky], qam = xq[km], qb = xq[F + ky], qbm = xq[F + km];
    const c32 fa0{0.5f * (qa.x + qam.x), 0.5f * (qa.y - qam.y)};
    const c32 fan{0.5f * (qa.y + qam.y), 0.5f * (qam.x - qa.x)};
    const c32 fb0{0.5f * (qb.x + qbm.x), 0.5f * (qb.y - qbm.y)};
    const c32 fbn{0.5f * (qb.y + qbm.y), 0.5f * (qbm.x - qb.x)};
    const c32 p0 = cconjmul(fa0, fb0), pn = cconjmul(fan, fbn);
    xq[2 * F + ky] = c32{p0.x - pn.y, p0.y + pn.x};
  }
  __syncthreads();
  if (kx == 0) {  // the A and B lanes of slot 0 both continue with q
    static_for<0, H>([&](auto KK) {
      constexpr int k = decltype(KK)::value;
      x[bitrev(k, H)] = xq[2 * F + k + (upper ? H : 0)];
    });
  }

  // ---- stage IC: inverse column FFT of P over the quad (4-lane DIF) -------
  c32 z[Q];
  dif2_split<F, +1>(x, upper);
  static_for<0, Q>([&](auto NN) {
    constexpr int nn = decltype(NN)::value;
    c32 y0 = x[bitrev(nn, H)];
    c32 y1 = x[bitrev(nn + Q, H)];
    const c32 t = sel ? csub(y0, y1) : cadd(y0, y1);
    z[nn] = sel ? twiddle<nn, H, +1>(t) : t;
  });
  fft_reg<Q, +1>(z);
  // lane (kx, sel, h) now holds R[4k + 2 sel + h][kx] at z[bitrev(k, Q)]
  static_for<0, Q>([&](auto KK) {
    constexpr int k = decltype(KK)::value;
    const int yy = 4 * k + 2 * sel + (upper ? 1 : 0);
    R[(size_t)yy * SR + kx] = z[bitrev(k, Q)];
  });
  __syncthreads();

  // ---- stage IR: inverse rows, rows (2j, 2j+1) packed as real/imag --------
  const int jpair = t >> 2;
  const int g = t & 3;
  {
    const int y1 = 2 * jpair, y2 = y1 + 1;
    static_for<0, Q>([&](auto MM) {
      constexpr int m = decltype(MM)::value;
      const int kxx = 4 * m + g;
      const int kk = kxx <= H ? kxx : F - kxx;
      const int slot = kk == H ? 0 : kk;
      c32 l1 = R[(size_t)y1 * SR + slot];
      c32 l2 = R[(size_t)y2 * SR + slot];
      c32 sv;
      if (kxx == 0)
        sv = c32{l1.x, l2.x};
      else if (kxx == H)
        sv = c32{l1.y, l2.y};
      else if (kxx < H)
        sv = c32{l1.x - l2.y, l1.y + l2.x};
      else
        sv = c32{l1.x + l2.y, l2.x - l1.y};
      z[m] = sv;
    });
  }
  fft_reg<Q, +1>(z);
  {
    const bool hi2 = g & 2;
    const bool hi1 = g & 1;
    // combine A (lanes g <-> g^2): half-length DIT butterflies
    static_for<0, Q>([&](auto II) {
      constexpr int i = decltype(II)::value;
      constexpr int r = bitrev(i, Q);
      const c32 own = z[r];
      const c32 v = hi2 ? twiddle<i, H, +1>(own) : own;
      const c32 recv = swap2(v);
      const c32 sv = hi2 ? c32{-v.x, -v.y} : v;
      z[r] = c32{recv.x + sv.x, recv.y + sv.y};
    });
    // combine B (lanes g <-> g^1); lanes with g & 2 carry the extra w_F^{F/4} = i
    static_for<0, Q>([&](auto II) {
      constexpr int i = decltype(II)::value;
      constexpr int r = bitrev(i, Q);
      const c32 own = z[r];
      const c32 t = twiddle<i, F, +1>(own);
      const c32 ti = hi2 ? c32{-t.y, t.x} : t;
      const c32 v = hi1 ? ti : own;
      const c32 recv = swap1(v);
      const c32 sv = hi1 ? c32{-v.x, -v.y} : v;
      z[r] = c32{recv.x + sv.x, recv.y + sv.y};
    });
  }
  __syncthreads();  // R fully consumed (plane aliases it)

  // ---- peak search over the fftshift-ed plane -----------------------------
  const float scale = 1.0f / (float)(4 * F * F);  // 4: the row split's factors of 2
  const int beta = (g == 0) ? 0 : (g == 1 ? 2 : (g == 2 ? 1 : 3));
  Best best;
  {
    const int i1 = (2 * jpair) ^ H;
    const int i2 = i1 + 1;
    // this lane's columns jj = (k + beta Q) ^ H increase with k; rows i1 < i2
    const int jj0 = (beta * Q) ^ H;
    float* gp = (a.planes && valid) ? a.planes + (size_t)win * F * F : nullptr;
    float zmax = -INFINITY;  // on the raw values: the scale is a power of two
    static_for<0, Q>([&](auto KK) {
      constexpr int k = decltype(KK)::value;
      const int jj = jj0 + k;
      const c32 val = z[bitrev(k, Q)];
      const float c1 = val.x * scale, c2 = val.y * scale;
      plane[i1 * SP + jj] = c1;
      plane[i2 * SP + jj] = c2;
      zmax = fmaxf(zmax, fmaxf(val.x, val.y));
      if (gp) {
        gp[i1 * F + jj] = c1;
        gp[i2 * F + jj] = c2;
      }
    });
    // first occurrence of the maximum: walk the lane's values backwards
    int first = 0x7fffffff;
    static_for<0, Q>([&](auto KK) {
      constexpr int k = Q - 1 - decltype(KK)::value;
      first = z[bitrev(k, Q)].y == zmax ? i2 * F + jj0 + k : first;
    });
    static_for<0, Q>([&](auto KK) {
      constexpr int k = Q - 1 - decltype(KK)::value;
      first = z[bitrev(k, Q)].x == zmax ? i1 * F + jj0 + k : first;
    });
    best = Best{zmax * scale, first};
  }
#pragma unroll
  for (int off = WIDTH / 2; off > 0; off >>= 1) {
    // opaque offset: the shuffle addresses are recomputed here instead of
    // being kept (spilled) from the window-sum reduction at the start
    int o_off = off;
    asm volatile("" : "+v"(o_off));
    Best o{__shfl_xor(best.v, o_off, 64), __shfl_xor(best.i, o_off, 64)};
    best = better(best, o);
  }
  if constexpr (LANES > 64) {
    if ((tid & 63) == 0) red_b[wave] = best;
  }
  __syncthreads();
  if constexpr (LANES > 64) {  // every lane of the window learns the first peak
    constexpr int WPW = LANES / 64;
    const int w0 = (tid / LANES) * WPW;
    best = red_b[w0];
    for (int k = 1; k < WPW; ++k) best = better(best, red_b[w0 + k]);
  }

  // ---- signal-to-noise ratio (CorrelationGPU.get_s2n_ratio,
  // gpu_process.py:1312-1365) on the same fftshift-ed plane ---------------
  float s2n = 0.f;
  if (EXTRA && a.s2n) {  // uniform over the launch
    const float c1 = best.v;
    const int ip1 = best.i / F, jp1 = best.i % F;
    float c2;
    if (a.s2n_method == 0) {
      // second peak: first-occurrence argmax outside the (2w+1)^2 block
      // around the first peak, the block wrapping around the plane edges
      // (fill_kernel through CuPy's wrapping fancy indices, gpu_misc.py:52-85)
      const int w = a.s2n_size;
      Best b2{-INFINITY, 0x7fffffff};
      for (int idx = t; idx < F * F; idx += LANES) {
        const int i = idx / F, j = idx % F;
        int di = i - ip1, dj = j - jp1;
        di = di < 0 ? -di : di;
        dj = dj < 0 ? -dj : dj;
        di = di < F - di ? di : F - di;
        dj = dj < F - dj ? dj : F - dj;
        if (di <= w && dj <= w) continue;
        b2 = better(b2, Best{plane[i * SP + j], idx});
      }
#pragma unroll
      for (int off = WIDTH / 2; off > 0; off >>= 1) {
        Best o{__shfl_xor(b2.v, off, 64), __shfl_xor(b2.i, off, 64)};
        b2 = better(b2, o);
      }
      if constexpr (LANES > 64) {
        if ((tid & 63) == 0) red_s[wave] = b2;
        __syncthreads();
        constexpr int WPW = LANES / 64;
        const int w0 = (tid / LANES) * WPW;
        b2 = red_s[w0];
        for (int k = 1; k < WPW; ++k) b2 = better(b2, red_s[w0 + k]);
      }
      c2 = b2.i == 0x7fffffff ? -INFINITY : b2.v;
    } else {
      // mean of the positive values (peak2energy), of those <= peak / 2 (peak2mean)
      const float lim = 0.5f * c1;
      float sum = 0.f;
      int cnt = 0;
      for (int idx = t; idx < F * F; idx += LANES) {
        const float val = plane[(idx / F) * SP + idx % F];
        if (val > 0.f && (a.s2n_method == 2 || val <= lim)) {
          sum += val;
          ++cnt;
        }
      }
#pragma unroll
      for (int off = WIDTH / 2; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off, 64);
        cnt += __shfl_xor(cnt, off, 64);
      }
      if constexpr (LANES > 64) {
        if ((tid & 63) == 0) red_s[wave] = Best{sum, cnt};
        __syncthreads();
        constexpr int WPW = LANES / 64;
        const int w0 = (tid / LANES) * WPW;
        sum = 0.f;
        cnt = 0;
        for (int k = 0; k < WPW; ++k) {
          sum += red_s[w0 + k].v;
          cnt += red_s[w0 + k].i;
        }
      }
      c2 = cnt > 0 ? sum / (float)cnt : NAN;
    }
    // :1329-1333 -- 0 where the first peak is not positive, inf where the
    // second is not, NaN when the mean is empty
    s2n = c1 > 0.f ? (c2 <= 0.f ? INFINITY : c1 / c2) : 0.f;
  }

  if (t == 0 && valid) {
    if (EXTRA && a.s2n) a.s2n[win] = s2n;
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
      uu += a.up[win];
      vv += a.vp[win];
    }
    a.u[win] = uu;
    a.v[win] = vv;
    if (a.i_peak) {
      a.i_peak[win] = ip;
      a.j_peak[win] = jp;
    }
  }
}

// Dynamic LDS above 64 KiB must be opted into; request exactly what a
// launch needs (gfx950 allows up to 160 KiB per workgroup).
template <int F, typename InT, bool BATCHED, bool EXTRA, bool HALF>
static hipError_t lds_attr(size_t lds) {
  static thread_local size_t attr_bytes = 0;
  if (lds > 65536 && lds > attr_bytes) {
    hipError_t e = hipFuncSetAttribute((const void*)corr_kernel<F, InT, BATCHED, EXTRA, HALF>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
      set_launch_detail("hipFuncSetAttribute(MaxDynamicSharedMemorySize)", e);
      return e;
    }
    attr_bytes = lds;
  }
  return hipSuccess;
}

template <int F, typename InT, bool EXTRA, bool HALF>
static hipError_t launch_fv(const CorrArgs& a, hipStream_t stream) {
  using L = Layout<F>;
  const int n_windows = a.m * a.n;
  const size_t region = correlate_lds_bytes(F, a.ss);
  const size_t lds = (region + 3 * F * sizeof(c32)) * L::WPB;  // + the slot-0 exchange
  const int blocks = (n_windows + L::WPB - 1) / L::WPB;
  (void)hipGetLastError();  // do not report an earlier, unrelated error
  hipError_t e = lds_attr<F, InT, false, EXTRA, HALF>(lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((corr_kernel<F, InT, false, EXTRA, HALF>), dim3(blocks), dim3(256), lds, stream, a,
                     CorrBatch{}, n_windows, (int)(region / sizeof(float)));
  e = hipGetLastError();
  if (e != hipSuccess) set_launch_detail("corr_kernel launch", e);
  return e;
}

template <int F, typename InT>
static hipError_t launch_f(const CorrArgs& a, hipStream_t stream) {
  const bool extra = a.s2n != nullptr || a.win_pad != 0;
  const bool half = 2 * a.ss <= F;
  if (extra) return half ? launch_fv<F, InT, true, true>(a, stream) : launch_fv<F, InT, true, false>(a, stream);
  return half ? launch_fv<F, InT, false, true>(a, stream) : launch_fv<F, InT, false, false>(a, stream);
}

template <int F, typename InT, bool HALF>
static hipError_t launch_fbv(const CorrBatch& b, int total_blocks, int max_ss, hipStream_t stream) {
  using L = Layout<F>;
  const size_t region = correlate_lds_bytes(F, max_ss);
  const size_t lds = (region + 3 * F * sizeof(c32)) * L::WPB;  // + the slot-0 exchange
  (void)hipGetLastError();
  hipError_t e = lds_attr<F, InT, true, true, HALF>(lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((corr_kernel<F, InT, true, true, HALF>), dim3(total_blocks), dim3(256), lds, stream, CorrArgs{},
                     b, 0, (int)(region / sizeof(float)));
  e = hipGetLastError();
  if (e != hipSuccess) set_launch_detail("corr_kernel (batched) launch", e);
  return e;
}

template <int F, typename InT>
static hipError_t launch_fb(const CorrBatch& b, int total_blocks, int max_ss, bool half, hipStream_t stream) {
  return half ? launch_fbv<F, InT, true>(b, total_blocks, max_ss, stream)
              : launch_fbv<F, InT, false>(b, total_blocks, max_ss, stream);
}

template <int F>
static hipError_t launch_tb(const CorrBatch& b, int in_dtype, int total_blocks, int max_ss, bool half,
                            hipStream_t stream) {
  switch (in_dtype) {
    case kU8: return launch_fb<F, uint8_t>(b, total_blocks, max_ss, half, stream);
    case kU16: return launch_fb<F, uint16_t>(b, total_blocks, max_ss, half, stream);
    case kF32: return launch_fb<F, float>(b, total_blocks, max_ss, half, stream);
  }
  return hipErrorInvalidValue;
}

template <int F>
static hipError_t launch_t(const CorrArgs& a, hipStream_t stream) {
  switch (a.in_dtype) {
    case kU8: return launch_f<F, uint8_t>(a, stream);
    case kU16: return launch_f<F, uint16_t>(a, stream);
    case kF32: return launch_f<F, float>(a, stream);
  }
  return hipErrorInvalidValue;
}

int correlate_windows_per_block(int fsize) { return 256 / (2 * fsize); }

hipError_t launch_correlate_batch(const CorrBatch& b, int fsize, int in_dtype, int total_blocks, int max_ss,
                                  bool half, hipStream_t stream) {
  if (total_blocks <= 0) return hipSuccess;
  switch (fsize) {
    case 16: return launch_tb<16>(b, in_dtype, total_blocks, max_ss, half, stream);
    case 32: return launch_tb<32>(b, in_dtype, total_blocks, max_ss, half, stream);
    case 64: return launch_tb<64>(b, in_dtype, total_blocks, max_ss, half, stream);
    case 128: return launch_tb<128>(b, in_dtype, total_blocks, max_ss, half, stream);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_correlate(const CorrArgs& a, hipStream_t stream) {
  if (a.win_pad < 0 || 2 * a.win_pad >= a.ss) {
    set_launch_detail("correlate: window_size / search_size mismatch", hipErrorInvalidValue);
    return hipErrorInvalidValue;
  }
  if (a.ss < 2 || a.ss > a.fsize || (a.fsize - a.ss) % 2 != 0) {
    set_launch_detail("correlate: search_size / fft_size mismatch", hipErrorInvalidValue);
    return hipErrorInvalidValue;
  }
  // F <= 32 with a search window <= F/2 and neither S2N nor a shrunk
  // interrogation window: the one-lane-per-line kernel
  if (correlate_line_enabled() && correlate_line_applies(a)) return launch_correlate_line(a, stream);
  // F = 64, 128 under the same conditions: the residue-split kernel
  if (correlate_split_enabled() && correlate_split_applies(a)) return launch_correlate_split(a, stream);
  if (a.fsize > kMaxSmallFft) return launch_correlate_big(a, stream);
  switch (a.fsize) {
    case 16: return launch_t<16>(a, stream);
    case 32: return launch_t<32>(a, stream);
    case 64: return launch_t<64>(a, stream);
    case 128: return launch_t<128>(a, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace piv
