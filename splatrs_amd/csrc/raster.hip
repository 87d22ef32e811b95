// Tile compositing, forward and backward (src/rasterize.rs:231-505).
//
// One 256-lane workgroup per 16x16 tile, one lane per pixel.  The tile's depth-sorted
// entries are staged through LDS in batches (record gather + per-tile pixel rectangle,
// TileGauss::new at src/rasterize.rs:149-182); every lane walks the batch with its own
// compositing state, exactly the per-pixel walk of the reference's entry-outer loops.
// The forward pass stops a workgroup when all its pixels are done.
//
// Backward: each entry's 11 per-pixel terms are reduced over the tile's pixels and
// written to the entry's slots (src/rasterize.rs:449-454); a second kernel sums each
// Gaussian's slots in entry order (src/rasterize.rs:480-503).  ORDERED=true sums the
// pixel terms sequentially in row-major order -- the reference's order, so slots and
// gradients are bitwise identical to it; ORDERED=false uses DPP row reductions.
#include "internal.hpp"

namespace splat {
namespace {

constexpr int kOrdBatch = 8;     // entries per batch (ordered backward)

struct TileInfo {
  uint32_t gx0, gy0, w, h;
};

__device__ __forceinline__ TileInfo tile_info(uint32_t tile, uint32_t tiles_x, uint32_t W,
                                              uint32_t H) {  // src/rasterize.rs:198-209
  TileInfo t;
  t.gx0 = (tile % tiles_x) * kTile;
  t.gy0 = (tile / tiles_x) * kTile;
  t.w = min(W - t.gx0, (uint32_t)kTile);
  t.h = min(H - t.gy0, (uint32_t)kTile);
  return t;
}

// Geometry of one tile of the (possibly batched) tile space: its pixel rectangle, the
// width of its view and the view's first pixel in the image buffers.
struct TileGeom {
  TileInfo ti;
  uint32_t width;
  size_t pix0;
};

template <bool BATCH>
__device__ __forceinline__ TileGeom tile_geom(uint32_t tile, uint32_t tiles_x, uint32_t W,
                                              uint32_t H, const BatchDev& bd) {
  TileGeom t;
  if (BATCH) {
    const uint32_t v = find_segment(bd.tile_base, bd.npairs, tile);
    t.ti = tile_info(tile - bd.tile_base[v], bd.tiles_x[v], bd.width[v], bd.height[v]);
    t.width = bd.width[v];
    t.pix0 = bd.pix_base[v];
  } else {
    t.ti = tile_info(tile, tiles_x, W, H);
    t.width = W;
    t.pix0 = 0;
  }
  return t;
}

// The workgroup's next tile: its block index, or (persistent launch) the next one handed
// out by a shared counter.  Every lane sees the same tile.
constexpr uint32_t kXcds = 8;  // accelerator complex dies, each with its own L2

// The die the calling wave runs on (hardware register XCC_ID).
__device__ __forceinline__ uint32_t xcc_id() {
  uint32_t x;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
  return x % kXcds;
}

// Persistent launches: die x takes, from its own counter, the tiles of its eighth of
// [0, T) in index order (neighbouring tiles share Gaussian records in that die's L2), then
// whatever the other eighths have left.  Counters only grow, so every range drains and the
// call returns T once all are taken.  Which die a wave reports only moves work between
// dies; every tile is still taken exactly once.
__device__ __forceinline__ uint32_t take_tile(uint32_t* counter, uint32_t T) {
  const uint32_t x = xcc_id();
  for (uint32_t k = 0; k < kXcds; k++) {
    const uint32_t r = (x + k) % kXcds;
    const uint32_t lo = (uint32_t)((uint64_t)T * r / kXcds);
    const uint32_t hi = (uint32_t)((uint64_t)T * (r + 1) / kXcds);
    const uint32_t t = atomicAdd(&counter[r], 1u);
    if (t < hi - lo) return lo + t;
  }
  return T;
}

// With a dispatch order (tiles by decreasing entry count) the t-th tile taken is order[t];
// a tile's results do not depend on when it runs.
template <bool PERSIST>
__device__ __forceinline__ uint32_t next_tile(uint32_t* counter, uint32_t* s_tile,
                                              const uint32_t* __restrict__ order,
                                              uint32_t num_tiles) {
  uint32_t t;
  if (!PERSIST) {
    t = blockIdx.x;
  } else {
    __syncthreads();  // the previous tile's LDS use is over
    if (threadIdx.x == 0) *s_tile = take_tile(counter, num_tiles);
    __syncthreads();
    t = *s_tile;
  }
  return order && t < num_tiles ? order[t] : t;
}

// Tile-local pixel range of one entry (the clip closure of src/rasterize.rs:158-164).
__device__ __forceinline__ void clip_range(float m, float half, uint32_t offset, uint32_t n,
                                           uint32_t* lo_out, uint32_t* hi_out) {
  const float lo = ceilf((m - half) - 0.5f) - (float)offset;
  const float hi = floorf((m + half) - 0.5f) - (float)offset + 1.0f;
  *lo_out = as_u32(rclamp(lo, 0.0f, (float)n));
  *hi_out = as_u32(rclamp(hi, 0.0f, (float)n));
}

// Packed tile-local rectangle x0 | x1 << 8 | y0 << 16 | y1 << 24 (all <= 16).
__device__ __forceinline__ uint32_t entry_rect(float4 r0, float4 r1, float4 r2, const TileInfo& t) {
  const float hx = r1.w, hy = r2.x;
  if (!(hx >= 0.0f)) return 0u;  // sigma_max <= 0 or det <= 0: empty
  uint32_t x0, x1, y0, y1;
  clip_range(r0.x, hx, t.gx0, t.w, &x0, &x1);
  clip_range(r0.y, hy, t.gy0, t.h, &y0, &y1);
  return x0 | (x1 << 8) | (y0 << 16) | (y1 << 24);
}

// The records as staged in LDS: conic A (r0.z) and C (r1.x) halved.  Scaling by 1/2
// commutes with every rounding of sigma = 0.5 (A dx dx + C dy dy) + B dx dy
// (src/rasterize.rs:235-236) as long as the products stay normal: a nonzero dx or dy is
// at least 2^-25 (an ulp of a coordinate in [0.5, 2^24)), so for conic entries above 2^-76
// (2D covariances below about 10^22 px^2) (A/2) dx dx + (C/2) dy dy is exactly half of
// A dx dx + C dy dy and the walks get the reference's sigma bit for bit with one multiply
// less.  Uses of the full A and C double them back (exact).
__device__ __forceinline__ float4 half_conic0(float4 r0) {
  r0.z *= 0.5f;
  return r0;
}
__device__ __forceinline__ float4 half_conic1(float4 r1) {
  r1.x *= 0.5f;
  return r1;
}

// Pixel of thread t in the 16x16 tile: wave w = t / 64 covers the 8x8 block
// (w % 2, w / 2); its 16-lane row r = l / 16 the 4x4 sub-block (r % 2, r / 2) of that
// block, and lane l the pixel (l % 4, (l / 4) % 4) of the sub-block.  Square blocks let the
// rectangle tests below cull in both directions, and the backward walks a list of entries
// per row (the DPP reductions are per row).
__device__ __forceinline__ void tile_pixel(uint32_t t, uint32_t* i, uint32_t* j) {
  const uint32_t w = t >> 6, r = (t >> 4) & 3u, q = t & 15u;
  *i = (w & 1u) * 8u + (r & 1u) * 4u + (q & 3u);
  *j = (w >> 1) * 8u + (r >> 1) * 4u + (q >> 2);
}

// Whether a packed rectangle has pixels in the 4x4 sub-block of row r of wave w.
__device__ __forceinline__ bool row_touches(uint32_t rc, uint32_t w, uint32_t r) {
  const uint32_t x0 = rc & 0xffu, x1 = (rc >> 8) & 0xffu, y0 = (rc >> 16) & 0xffu, y1 = rc >> 24;
  const uint32_t bx = (w & 1u) * 8u + (r & 1u) * 4u, by = (w >> 1) * 8u + (r >> 1) * 4u;
  return (x0 < x1) & (y0 < y1) & (x0 < bx + 4u) & (x1 > bx) & (y0 < by + 4u) & (y1 > by);
}

// Whether a packed rectangle has pixels in the 8x8 block of wave w.
__device__ __forceinline__ bool wave_touches(uint32_t rc, uint32_t w) {
  const uint32_t x0 = rc & 0xffu, x1 = (rc >> 8) & 0xffu, y0 = (rc >> 16) & 0xffu, y1 = rc >> 24;
  const uint32_t bx = (w & 1u) * 8u, by = (w >> 1) * 8u;
  return (x0 < x1) & (y0 < y1) & (x0 < bx + 8u) & (x1 > bx) & (y0 < by + 8u) & (y1 > by);
}

__device__ __forceinline__ bool in_rect(uint32_t rc, uint32_t i, uint32_t j) {
  // bitwise, not short-circuit: no divergent branches in the compositing loops
  return (i >= (rc & 0xffu)) & (i < ((rc >> 8) & 0xffu)) & (j >= ((rc >> 16) & 0xffu)) &
         (j < (rc >> 24));
}

// evaluate, src/rasterize.rs:231-244.
__device__ __forceinline__ bool evaluate(float4 r0, float4 r1, float px, float py,
                                         const uint64_t* tab, float* gauss, float* alpha) {
  const float dx = px - r0.x;
  const float dy = py - r0.y;
  const float sigma = 0.5f * (r0.z * dx * dx + r1.x * dy * dy) + r0.w * dx * dy;
  if (sigma < 0.0f || sigma > r1.z) return false;
  const float G = expf_tab(-sigma, tab);
  const float og = r1.y * G;
  if (og < kAlphaMin) return false;
  *gauss = G;
  *alpha = rmin(og, kAlphaMax);
  return true;
}

// Pair counters of the counting instantiations (COUNT = true), one 64-bit word each.
// Forward: (entry, wave) visits, visits in which some lane evaluates the Gaussian,
// evaluated pairs (sigma in [0, sigma_max]: the exponential is computed), passing pairs
// (alpha >= 1/255 on a pixel still compositing) and composited pairs.  Backward: visits,
// active visits, evaluated and passing pairs of the reverse walk.
enum PairCounter : int {
  kFwdVisits, kFwdActive, kFwdEvaluated, kFwdPassing, kFwdComposited,
  kBwdVisits, kBwdActive, kBwdEvaluated, kBwdPassing, kNumPairCounters
};

// Adds a wave's counts to the global counters (lane 0; the counts are wave-uniform).
__device__ __forceinline__ void flush_counts(unsigned long long* pairs, int first,
                                             const uint32_t* c, int n) {
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < n; q++)
      if (c[q]) atomicAdd(&pairs[first + q], (unsigned long long)c[q]);
}

// B: entries per LDS batch; BATCH: tiles of many views (BatchDev); PERSIST: a fixed grid
// that takes tiles from `counter` until `num_tiles` are done; COUNT: also count pairs
// into `pairs` (PairCounter).
template <int B, bool BATCH, bool PERSIST, bool COUNT = false>
__global__ void __launch_bounds__(256) raster_fwd_kernel(
    const float4* __restrict__ rec, const uint32_t* __restrict__ tile_start,
    const uint32_t* __restrict__ order,
    const uint32_t* __restrict__ sorted_gauss, uint32_t num_tiles, uint32_t tiles_x, uint32_t W,
    uint32_t H, float bg0, float bg1, float bg2, float* __restrict__ image,
    float* __restrict__ final_t, uint32_t* __restrict__ n_contrib, BatchDev bd,
    uint32_t* __restrict__ counter, unsigned long long* __restrict__ pairs) {
  static_assert(B <= 256, "each of the 256 lanes stages at most one entry of a batch");
  __shared__ uint64_t tab[32];
  __shared__ float4 s0[B], s1[B], s2[B];
  __shared__ uint32_t srect[B];
  __shared__ uint32_t s_tile;
  if (threadIdx.x < 32) tab[threadIdx.x] = kExp2Table[threadIdx.x];
  for (;;) {
  const uint32_t tile = next_tile<PERSIST>(counter, &s_tile, order, num_tiles);
  if (tile >= num_tiles) break;
  const TileGeom tg = tile_geom<BATCH>(tile, tiles_x, W, H, bd);
  const TileInfo ti = tg.ti;
  uint32_t i, j;
  tile_pixel(threadIdx.x, &i, &j);
  const bool valid = i < ti.w && j < ti.h;
  const float px = (float)(ti.gx0 + i) + 0.5f;
  const float py = (float)(ti.gy0 + j) + 0.5f;
  float T = 1.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
  uint32_t nc = 0;
  bool done = !valid;
  uint32_t pcount[5] = {0, 0, 0, 0, 0};
  const uint32_t start = tile_start[tile], end = tile_start[tile + 1];
  for (uint32_t b0 = start; b0 < end; b0 += B) {
    if (__syncthreads_and(done ? 1 : 0)) break;
    const uint32_t cnt = min((uint32_t)B, end - b0);
    if (threadIdx.x < cnt) {
      const uint32_t g = sorted_gauss[b0 + threadIdx.x];
      const float4 a = rec[4 * g], b = rec[4 * g + 1], c = rec[4 * g + 2];
      s0[threadIdx.x] = half_conic0(a);
      s1[threadIdx.x] = half_conic1(b);
      s2[threadIdx.x] = c;
      srect[threadIdx.x] = entry_rect(a, b, c, ti);
    }
    __syncthreads();
    // Predicated form of the per-pixel walk (src/rasterize.rs:262-290): every decision of
    // evaluate() and the T_MIN test is a lane predicate; only the exp is skipped when no
    // lane of the wave needs it.  Each wave (an 8x8 pixel block) walks, in order, only the
    // entries whose pixel rectangle reaches its block (a ballot over the batch).
    for (uint32_t c0k = 0; c0k < cnt; c0k += 64) {
    const uint32_t kl = c0k + (threadIdx.x & 63);
    uint64_t todo = __ballot(kl < cnt && wave_touches(srect[min(kl, cnt - 1)], threadIdx.x >> 6));
    while (todo) {
      const uint32_t k = c0k + (uint32_t)__builtin_ctzll(todo);
      todo &= todo - 1;
      const float4 a = s0[k], b = s1[k];
      const float dx = px - a.x;
      const float dy = py - a.y;
      const float sigma = (a.z * dx * dx + b.x * dy * dy) + a.w * dx * dy;  // halved A, C
      const bool pass1 = (!done) & in_rect(srect[k], i, j) & !((sigma < 0.0f) | (sigma > b.z));
      if (COUNT) {
        const uint64_t b1 = __ballot(pass1);
        pcount[0]++;
        if (!b1) continue;
        pcount[1]++;
        pcount[2] += (uint32_t)__popcll(b1);
      } else if (!__any(pass1)) {
        continue;
      }
      const float G = expf_walk_tab(-sigma, tab);  // used where 0 <= sigma <= sigma_max
      const float og = b.y * G;
      const bool pass2 = pass1 & !(og < kAlphaMin);
      const float alpha = rmin(og, kAlphaMax);
      const float test_t = T * (1.0f - alpha);
      const bool stop = test_t < kTMin;
      done = done | (pass2 & stop);
      const bool comp = pass2 & !stop;
      if (COUNT) {
        pcount[3] += (uint32_t)__popcll(__ballot(pass2));
        pcount[4] += (uint32_t)__popcll(__ballot(comp));
      }
      const float4 c = s2[k];
      const float w = alpha * T;
      c0 = comp ? c0 + w * c.y : c0;
      c1 = comp ? c1 + w * c.z : c1;
      c2 = comp ? c2 + w * c.w : c2;
      T = comp ? test_t : T;
      nc = comp ? b0 - start + k + 1 : nc;
    }
    }
  }
  if (valid) {
    const size_t pix = tg.pix0 + (size_t)(ti.gy0 + j) * tg.width + ti.gx0 + i;
    image[pix * 3 + 0] = c0 + T * bg0;
    image[pix * 3 + 1] = c1 + T * bg1;
    image[pix * 3 + 2] = c2 + T * bg2;
    final_t[pix] = T;
    n_contrib[pix] = nc;
  }
  if (COUNT) flush_counts(pairs, kFwdVisits, pcount, 5);
  if (!PERSIST) break;
  }
}

// Per-pixel backward state (PixelBack, src/rasterize.rs:345-350).
struct PixBack {
  float t, acc0, acc1, acc2, last_alpha, lc0, lc1, lc2;
  float d0, d1, d2, bg_dot, t_final;
  uint32_t last;
};

// Terms of one (entry, pixel) pair, src/rasterize.rs:404-445, in predicated form: `act`
// says the pixel is inside the entry's rectangle and before its last contributor.
// Returns whether any lane of the wave contributes; v[] holds the slot terms (zero for
// lanes that do not), and the pixel state advances only on contributing lanes.  The
// reference's expressions and correctly rounded divisions: the ordered backward is bitwise.
__device__ __forceinline__ bool pair_terms(PixBack& s, bool act, float4 r0, float4 r1, float4 r2,
                                           float px, float py, const uint64_t* tab,
                                           float v[kSlots]) {
#pragma unroll
  for (int q = 0; q < kSlots; q++) v[q] = 0.0f;
  const float dx = px - r0.x;
  const float dy = py - r0.y;
  const float sigma = (r0.z * dx * dx + r1.x * dy * dy) + r0.w * dx * dy;  // halved A, C
  const bool p1 = act & !((sigma < 0.0f) | (sigma > r1.z));
  if (!__any(p1)) return false;
  const float G = expf_walk_tab(-sigma, tab);  // used where 0 <= sigma <= sigma_max
  const float o = r1.y;
  const float og = o * G;
  const bool p2 = p1 & !(og < kAlphaMin);
  const float alpha = rmin(og, kAlphaMax);
  const float t = s.t / (1.0f - alpha);
  const float a0 = s.last_alpha * s.lc0 + (1.0f - s.last_alpha) * s.acc0;
  const float a1 = s.last_alpha * s.lc1 + (1.0f - s.last_alpha) * s.acc1;
  const float a2 = s.last_alpha * s.lc2 + (1.0f - s.last_alpha) * s.acc2;
  float dalpha = 0.0f;
  dalpha += (r2.y - a0) * s.d0;
  dalpha += (r2.z - a1) * s.d1;
  dalpha += (r2.w - a2) * s.d2;
  dalpha *= t;
  dalpha += (-s.t_final / (1.0f - alpha)) * s.bg_dot;
  const bool p3 = p2 & (og < kAlphaMax);
  const float dgauss = o * dalpha;
  const float dsigma = -dgauss * G;
  const float ddx = dsigma * ((2.0f * r0.z) * dx + r0.w * dy);
  const float ddy = dsigma * (r0.w * dx + (2.0f * r1.x) * dy);
  v[6] = p2 ? alpha * t * s.d0 : 0.0f;
  v[7] = p2 ? alpha * t * s.d1 : 0.0f;
  v[8] = p2 ? alpha * t * s.d2 : 0.0f;
  v[5] = p3 ? G * dalpha : 0.0f;
  v[0] = p3 ? -ddx : 0.0f;
  v[1] = p3 ? -ddy : 0.0f;
  v[2] = p3 ? dsigma * 0.5f * dx * dx : 0.0f;
  v[3] = p3 ? dsigma * dx * dy : 0.0f;
  v[4] = p3 ? dsigma * 0.5f * dy * dy : 0.0f;
  v[9] = p3 ? fabsf(ddx) : 0.0f;
  v[10] = p3 ? fabsf(ddy) : 0.0f;
  if (p2) {
    s.t = t;
    s.acc0 = a0;
    s.acc1 = a1;
    s.acc2 = a2;
    s.lc0 = r2.y;
    s.lc1 = r2.z;
    s.lc2 = r2.w;
    s.last_alpha = alpha;
  }
  return true;
}

// pair_terms in the tree-reduced form, rearranged for fewer instructions.  Equal up to
// rounding; the decisions (sigma range, alpha_min, alpha_max) are the exact ones, so the
// walk visits the same pairs as the forward:
//  * one hardware reciprocal y (1 ulp) of d = 1 - alpha serves both divisions by d.  The
//    transmittance, which the walk divides again at every contributor, takes one residual
//    correction, t = q + (s.t - q d) y with q = s.t y and the residual exact in an fma: the
//    error before the final rounding is O(ulp^2), so t is the correctly rounded quotient
//    (the reference's) but for inputs within 2^-45 of a rounding midpoint, and no bias
//    of the reciprocal builds up along the walk.  The background term's T_final / d takes
//    the same correction; with a black background (BG = false) that term is zero and is
//    left out;
//  * the accumulated colour behind the next contributor is updated eagerly,
//    A <- A + alpha (c - A), the reference's last_alpha * last_color + (1 - last_alpha) * acc
//    (src/rasterize.rs:420-428) evaluated one entry earlier;
//  * with q = dL/dsigma the slots 0-4 become the per-pixel products q dx, q dy, q dx^2,
//    q dx dy, q dy^2; the entry's conic turns the sums into dL/dmean and 0.5 scales the
//    diagonal conic slots once per entry, when the batch's slots are written;
//  * products with a zero factor stand in for the per-slot selects.
// The pixel state uses acc0..2 for A; lc and last_alpha are unused.
// COUNT: cnt[0..3] += visit, active visit, evaluated and passing pairs.
template <bool ABS, bool COUNT = false>
__device__ __forceinline__ bool pair_terms_fast(PixBack& s, bool act, float4 r0, float4 r1,
                                                float4 r2, float px, float py,
                                                const uint64_t* tab, bool bg, float v[kSlots],
                                                uint32_t* cnt = nullptr) {
  const float dx = px - r0.x;
  const float dy = py - r0.y;
  const float sigma = (r0.z * dx * dx + r1.x * dy * dy) + r0.w * dx * dy;  // halved A, C
  const bool p1 = act & !((sigma < 0.0f) | (sigma > r1.z));
  if (COUNT) {
    const uint64_t b1 = __ballot(p1);
    cnt[0]++;
    if (!b1) return false;
    cnt[1]++;
    cnt[2] += (uint32_t)__popcll(b1);
  } else if (!__any(p1)) {
    return false;
  }
  const float G = expf_walk_tab(-sigma, tab);  // used where 0 <= sigma <= sigma_max
  const float o = r1.y;
  const float og = o * G;
  const bool p2 = p1 & !(og < kAlphaMin);
  const bool p3 = p2 & (og < kAlphaMax);
  if (COUNT) cnt[3] += (uint32_t)__popcll(__ballot(p2));
  const float alpha = rmin(og, kAlphaMax);
  const float den = 1.0f - alpha;
  const float inv = __builtin_amdgcn_rcpf(den);
  const float q0 = s.t * inv;
  const float t = fmaf(fmaf(-q0, den, s.t), inv, q0);
  const float e0 = r2.y - s.acc0, e1 = r2.z - s.acc1, e2 = r2.w - s.acc2;
  float dalpha = e0 * s.d0;
  dalpha = fmaf(e1, s.d1, dalpha);
  dalpha = fmaf(e2, s.d2, dalpha);
  if (bg) {  // uniform: a kernel argument
    const float f0 = s.t_final * inv;
    const float tf = fmaf(fmaf(-f0, den, s.t_final), inv, f0);
    dalpha = fmaf(dalpha, t, -tf * s.bg_dot);
  } else {
    dalpha *= t;
  }
  const float q = p3 ? -(o * dalpha) * G : 0.0f;
  const float w = p2 ? alpha * t : 0.0f;
  const float qdx = q * dx, qdy = q * dy;
  v[0] = qdx;
  v[1] = qdy;
  v[2] = qdx * dx;
  v[3] = qdx * dy;
  v[4] = qdy * dy;
  v[5] = p3 ? G * dalpha : 0.0f;
  v[6] = w * s.d0;
  v[7] = w * s.d1;
  v[8] = w * s.d2;
  if (ABS) {
    v[9] = fabsf(fmaf(2.0f * r0.z, qdx, r0.w * qdy));
    v[10] = fabsf(fmaf(r0.w, qdx, (2.0f * r1.x) * qdy));
  } else {
    v[9] = v[10] = 0.0f;
  }
  s.t = p2 ? t : s.t;
  s.acc0 = p2 ? fmaf(alpha, e0, s.acc0) : s.acc0;
  s.acc1 = p2 ? fmaf(alpha, e1, s.acc1) : s.acc1;
  s.acc2 = p2 ? fmaf(alpha, e2, s.acc2) : s.acc2;
  return true;
}

// The value of lane (l ^ m) within each 16-lane row, by a DPP control whose partner map is
// that xor: row_ror:8 (m = 8), row_half_mirror (7), quad_perm [2,3,0,1] (2) and
// [1,0,3,2] (1).  The compiler folds the move into the add that consumes it.
template <int CTRL>
__device__ __forceinline__ float dpp_row(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, true));
}
constexpr int kDppXor8 = 0x128, kDppXor7 = 0x141, kDppXor2 = 0x4e, kDppXor1 = 0xb1;

// Reduce-scatter of 8 values over each 16-lane row: afterwards lanes 2m and 2m+1 of the row
// hold the row's sum of v[m].  Each step pairs a lane with lane ^ {8, 7, 2, 1}; the two
// partners keep complementary halves of their values (chosen by bit 3, 2, 1 of the lane)
// and add the half the other sends.  The four masks span all 16 lanes, so every sum covers
// the row; the order of the additions is fixed.
//
// The first two steps select by lane bits 3 and 2, which are exactly the DPP banks (lanes
// 4k .. 4k + 3 of a row): two bank-masked DPP adds, each writing half of the row, give
// x + x(partner) on one half and y + y(partner) on the other.  The compiler does not form
// masked DPP adds, so they are written out; the s_nops give the DPP source reads the wait
// states they need after the VALU writes around the block.
__device__ __forceinline__ float dpp_halves_xor8(float x, float y) {  // banks 0-1: x, 2-3: y
  float out;
  asm volatile(
      "s_nop 1\n\t"
      "v_add_f32_dpp %0, %1, %1 row_ror:8 row_mask:0xf bank_mask:0x3\n\t"
      "v_add_f32_dpp %0, %2, %2 row_ror:8 row_mask:0xf bank_mask:0xc\n\t"
      "s_nop 1"
      : "=&v"(out)
      : "v"(x), "v"(y));
  return out;
}
__device__ __forceinline__ float dpp_halves_xor7(float x, float y) {  // banks 0, 2: x; 1, 3: y
  float out;
  asm volatile(
      "s_nop 1\n\t"
      "v_add_f32_dpp %0, %1, %1 row_half_mirror row_mask:0xf bank_mask:0x5\n\t"
      "v_add_f32_dpp %0, %2, %2 row_half_mirror row_mask:0xf bank_mask:0xa\n\t"
      "s_nop 1"
      : "=&v"(out)
      : "v"(x), "v"(y));
  return out;
}

__device__ __forceinline__ float row_scatter8(const float v[8], uint32_t lane) {
  const bool b1 = (lane & 2u) != 0;
  float a[4], b[2];
#pragma unroll
  for (int q = 0; q < 4; q++) a[q] = dpp_halves_xor8(v[q], v[q + 4]);
#pragma unroll
  for (int q = 0; q < 2; q++) b[q] = dpp_halves_xor7(a[q], a[q + 2]);
  const float c = (b1 ? b[1] : b[0]) + dpp_row<kDppXor2>(b1 ? b[0] : b[1]);
  return c + dpp_row<kDppXor1>(c);
}

// Sum over each 16-lane row by a rotate butterfly: every lane ends with its row's sum.
__device__ __forceinline__ float row_allsum16(float v) {
  v += dpp_row<0x121>(v);  // row_ror:1
  v += dpp_row<0x122>(v);  // row_ror:2
  v += dpp_row<0x124>(v);  // row_ror:4
  v += dpp_row<0x128>(v);  // row_ror:8
  return v;
}

// ORDERED: per-pixel terms staged in LDS and summed sequentially in row-major order.
// Otherwise: 16-lane row sums by DPP, 16 row partials per entry summed in a fixed order.
// ABS: also accumulate the AbsGS slots (9, 10); otherwise they are written as zero.
// BF: entries per LDS batch of the tree-reduced form (the ordered form uses kOrdBatch).
// ROWS (tree-reduced form): every 16-lane row walks its own entry list; otherwise the wave
// walks the entries that reach any of its rows (the same sums either way).  COUNT
// (tree-reduced form only): also count pairs into `pairs` (PairCounter).
template <bool ORDERED, bool ABS, int BF, bool BATCH, bool PERSIST, bool ROWS,
          bool COUNT = false>
__global__ void __launch_bounds__(256) raster_bwd_kernel(
    const float4* __restrict__ rec, const uint32_t* __restrict__ tile_start,
    const uint32_t* __restrict__ order,
    const uint32_t* __restrict__ sorted_gauss, const uint32_t* __restrict__ sorted_entry,
    uint32_t num_tiles, uint32_t tiles_x, uint32_t W, uint32_t H, float bg0, float bg1,
    float bg2, const float* __restrict__ final_t, const uint32_t* __restrict__ n_contrib,
    const float* __restrict__ dimage, float* __restrict__ slots, BatchDev bd,
    uint32_t* __restrict__ counter, unsigned long long* __restrict__ pairs) {
  static_assert(!(COUNT && ORDERED), "pairs are counted in the tree-reduced form");
  constexpr int B = ORDERED ? kOrdBatch : BF;
  __shared__ uint64_t tab[32];
  __shared__ float4 s0[B], s1[B], s2[B];
  __shared__ uint32_t srect[B];
  __shared__ uint32_t smax_last;
  __shared__ uint32_t s_tile;
  // per (wave, row): the batch entries the row walks, back to front
  __shared__ uint8_t slist[ORDERED ? 1 : 16][ORDERED ? 1 : B];
  // words per entry in `slots`: the absolute sums (slots 9, 10) only when asked for
  constexpr uint32_t SS = ABS ? kSlots : 9;
  // ORDERED: per-pixel terms of the batch; otherwise per-row partial sums of the SS slots
  // (18 KB at B = 32 without the absolute sums; the workgroup's LDS is then about 20.4 KB,
  // seven workgroups per CU, as many as its 72 VGPRs allow).
  __shared__ float red[ORDERED ? B * kSlots * 256 : B * SS * 16];
  // whether the background term (T_final bg . dL) can be non-zero for a finite dL
  const bool bg_any = (bg0 != 0.0f) | (bg1 != 0.0f) | (bg2 != 0.0f);
  if (threadIdx.x < 32) tab[threadIdx.x] = kExp2Table[threadIdx.x];
  for (;;) {
  const uint32_t tile = next_tile<PERSIST>(counter, &s_tile, order, num_tiles);
  if (tile >= num_tiles) break;
  if (threadIdx.x == 0) smax_last = 0;
  const TileGeom tg = tile_geom<BATCH>(tile, tiles_x, W, H, bd);
  const TileInfo ti = tg.ti;
  uint32_t i, j;
  tile_pixel(threadIdx.x, &i, &j);
  const bool valid = i < ti.w && j < ti.h;
  const float px = (float)(ti.gx0 + i) + 0.5f;
  const float py = (float)(ti.gy0 + j) + 0.5f;
  PixBack s;
  s.acc0 = s.acc1 = s.acc2 = 0.0f;
  s.lc0 = s.lc1 = s.lc2 = 0.0f;
  s.last_alpha = 0.0f;
  s.t = 1.0f;
  s.t_final = 1.0f;
  s.d0 = s.d1 = s.d2 = 0.0f;
  s.bg_dot = 0.0f;
  s.last = 0;
  if (valid) {
    const size_t pix = tg.pix0 + (size_t)(ti.gy0 + j) * tg.width + ti.gx0 + i;
    s.d0 = dimage[pix * 3 + 0];
    s.d1 = dimage[pix * 3 + 1];
    s.d2 = dimage[pix * 3 + 2];
    s.bg_dot = bg0 * s.d0 + bg1 * s.d1 + bg2 * s.d2;
    s.last = n_contrib[pix];
    s.t_final = final_t[pix];
    s.t = s.t_final;
  }
  const uint32_t start = tile_start[tile], end = tile_start[tile + 1];
  const uint32_t count = end - start;
  uint32_t pcount[4] = {0, 0, 0, 0};
  __syncthreads();
  if (valid && s.last > 0) atomicMax(&smax_last, s.last);
  __syncthreads();
  const uint32_t kmax = min(smax_last, count);
  // Entries no pixel reaches get zero slots.
  for (uint32_t q = kmax * SS + threadIdx.x; q < count * SS; q += blockDim.x) {
    const uint32_t k = q / SS;
    slots[(size_t)sorted_entry[start + k] * SS + q % SS] = 0.0f;
  }
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, row = lane >> 4;
  // the last contributor of any pixel of each 16-lane row of this wave
  uint32_t row_last[4];
  {
    uint32_t m = valid ? s.last : 0u;
#pragma unroll
    for (int d = 8; d >= 1; d >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d, 64));
#pragma unroll
    for (int r = 0; r < 4; r++) row_last[r] = (uint32_t)__builtin_amdgcn_readlane((int)m, 16 * r);
  }
  // Batches walk [0, kmax) from the back.
  for (int32_t hi = (int32_t)kmax; hi > 0; hi -= B) {
    const int32_t lo = max(hi - B, 0);
    const uint32_t cnt = (uint32_t)(hi - lo);
    __syncthreads();
    if (threadIdx.x < cnt) {
      const uint32_t g = sorted_gauss[start + lo + threadIdx.x];
      const float4 a = rec[4 * g], b = rec[4 * g + 1], c = rec[4 * g + 2];
      s0[threadIdx.x] = half_conic0(a);
      s1[threadIdx.x] = half_conic1(b);
      s2[threadIdx.x] = c;
      srect[threadIdx.x] = entry_rect(a, b, c, ti);
    }
    if (!ORDERED)  // partials of entries a wave skips stay zero
      for (uint32_t q = threadIdx.x; q < cnt * SS * 16; q += blockDim.x) red[q] = 0.0f;
    __syncthreads();
    if (ORDERED) {
      for (int32_t kk = (int32_t)cnt - 1; kk >= 0; kk--) {
        const uint32_t k = lo + kk;
        float v[kSlots];
        const bool act = valid & (k < s.last) & in_rect(srect[kk], i, j);
        pair_terms(s, act, s0[kk], s1[kk], s2[kk], px, py, tab, v);
#pragma unroll
        for (int q = 0; q < kSlots; q++) red[(kk * kSlots + q) * 256 + j * kTile + i] = v[q];
      }
    } else {
      // Each 16-lane row walks, back to front, only the batch entries that reach its 4x4
      // sub-block and precede the row's last contributor; the four rows of a wave walk
      // their own lists side by side (B <= 64: one ballot per row covers the batch).
      // Sums over each 16-lane row: slots 0-7 by one reduce-scatter (even lane 2m holds
      // slot m), slot 8 (and the AbsGS slots 9, 10) by a butterfly read from lane 15 (13,
      // 11); one LDS store per writer lane.
      const uint32_t l16 = lane & 15u;
      const bool even = (l16 & 1u) == 0;
      const bool writer = even | (l16 == 15) | (ABS & ((l16 == 13) | (l16 == 11)));
      const uint32_t wslot = even ? (l16 >> 1) : (l16 == 15 ? 8u : (l16 == 13 ? 9u : 10u));
      const uint32_t woff = wslot * 16 + wave * 4 + row;
      // the rows' lists: entry kk sits at position popcount(mask >> (kk + 1))
      uint8_t* my_list = slist[wave * 4 + row];
      const uint32_t kl = min(lane, cnt - 1);
      uint32_t steps = 0, my_len = 0;
      uint64_t wave_todo = 0;  // !ROWS: the wave's entries, walked from the top bit down
      if (ROWS) {
        uint32_t len[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const uint64_t m = __ballot(lane < cnt && lo + lane < row_last[r] &&
                                      row_touches(srect[kl], wave, (uint32_t)r));
          len[r] = (uint32_t)__popcll(m);
          if ((m >> lane) & 1ull)
            slist[wave * 4 + r][__popcll(m >> lane) - 1] = (uint8_t)lane;
        }
        steps = max(max(len[0], len[1]), max(len[2], len[3]));
        my_len = row == 0 ? len[0] : (row == 1 ? len[1] : (row == 2 ? len[2] : len[3]));
        // the lists are written and read by this wave only
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      } else {
        const uint32_t wave_last =
            max(max(row_last[0], row_last[1]), max(row_last[2], row_last[3]));
        wave_todo = __ballot(lane < cnt && lo + lane < wave_last &&
                             wave_touches(srect[kl], wave));
        steps = (uint32_t)__popcll(wave_todo);
        my_len = steps;
      }
      for (uint32_t t = 0; t < steps; t++) {
        const bool has = t < my_len;
        uint32_t kk;
        if (ROWS) {
          kk = has ? (uint32_t)my_list[t] : 0u;
        } else {
          kk = 63u - (uint32_t)__builtin_clzll(wave_todo);
          wave_todo &= ~(1ull << kk);
        }
        const uint32_t k = lo + kk;
        float v[kSlots];
        const bool act = has & valid & (k < s.last) & in_rect(srect[kk], i, j);
        if (pair_terms_fast<ABS, COUNT>(s, act, s0[kk], s1[kk], s2[kk], px, py, tab, bg_any, v,
                                        pcount)) {
          float r = row_scatter8(v, lane);
          const float e8 = row_allsum16(v[8]);
          r = l16 == 15 ? e8 : r;
          if (ABS) {
            const float e9 = row_allsum16(v[9]);
            const float e10 = row_allsum16(v[10]);
            r = l16 == 13 ? e9 : r;
            r = l16 == 11 ? e10 : r;
          }
          if (writer & has) red[kk * SS * 16 + woff] = r;
        }
      }
    }
    __syncthreads();
    // Reduce the batch and write its slots.
    for (uint32_t q = threadIdx.x; q < cnt * SS; q += blockDim.x) {
      const uint32_t kk = q / SS, slot = q % SS;
      float acc = 0.0f;
      if (ORDERED) {
        const float* p = &red[(kk * kSlots + slot) * 256];
        for (int t = 0; t < 256; t++) acc += p[t];
      } else {
        // pair_terms_fast's products: slots 0, 1 combine both position sums with the conic
        const auto part_sum = [&](uint32_t sl) {
          const float* p = &red[(kk * SS + sl) * 16];
          float a = 0.0f;
#pragma unroll
          for (int t = 0; t < 16; t++) a += p[t];
          return a;
        };
        if (slot < 2) {
          const float sx = part_sum(0), sy = part_sum(1);
          const float4 a = s0[kk];
          acc = slot == 0 ? -fmaf(2.0f * a.z, sx, a.w * sy)
                          : -fmaf(a.w, sx, (2.0f * s1[kk].x) * sy);
        } else if (ABS || slot < 9) {
          acc = part_sum(slot);
          if (slot == 2 || slot == 4) acc *= 0.5f;
        }
      }
      slots[(size_t)sorted_entry[start + lo + kk] * SS + slot] = acc;
    }
  }
  if (COUNT) flush_counts(pairs, kBwdVisits, pcount, 4);
  if (!PERSIST) break;
  }
}

// Sum each row's slots over its entries in entry order (src/rasterize.rs:480-503).  One
// wave per 64 / NS consecutive rows, one lane per (row, slot): the rows' entries form one
// span, staged through LDS in windows with coalesced loads, and each lane adds its slot of
// its row's entries in order, so every sum keeps the reference's sequential order.  NS is
// the record width the backward wrote: 9, or 11 with the absolute sums (slots 9 and 10,
// into abs_grads).  A window of 128 entries holds the span of a wave's rows in the dense
// batched workload (about 100 entries) and keeps the workgroup's LDS at 4.6 KB (9 slots),
// so that LDS does not limit how many waves a CU holds (256 entries: 9.2 KB, about half).
constexpr uint32_t kReduceWindow = 128;  // entries staged per window

template <int NS>
__global__ void __launch_bounds__(64) reduce_slots_kernel(const uint32_t* __restrict__ entry_offset,
                                                          const float* __restrict__ slots,
                                                          uint32_t n,
                                                          splat_proj_grad* __restrict__ grads,
                                                          float* __restrict__ abs_grads,
                                                          const StepStatus* __restrict__ status) {
  constexpr uint32_t kRows = 64 / NS;
  constexpr uint32_t kLoads = 16;  // loads per lane in flight before their LDS stores
  __shared__ float sh[kReduceWindow * NS];
  if (status && status->abort) return;  // uniform
  const uint32_t lane = threadIdx.x;
  const uint32_t r0 = blockIdx.x * kRows, rl = lane / NS, k = lane % NS;
  const uint32_t row = r0 + rl;
  const bool live = rl < kRows && row < n;
  const uint32_t span0 = entry_offset[r0], span1 = entry_offset[min(r0 + kRows, n)];
  const uint32_t e0 = live ? entry_offset[row] : 0u, e1 = live ? entry_offset[row + 1] : 0u;
  float sum = 0.0f;
  for (uint32_t c0 = span0; c0 < span1; c0 += kReduceWindow) {
    const uint32_t words = (min(c0 + kReduceWindow, span1) - c0) * NS;
    const float* src = slots + (size_t)c0 * NS;
    __syncthreads();
    for (uint32_t q0 = 0; q0 < words; q0 += 64 * kLoads) {
      float t[kLoads];
#pragma unroll
      for (uint32_t u = 0; u < kLoads; u++) {
        const uint32_t q = q0 + u * 64 + lane;
        t[u] = src[min(q, words - 1)];  // in the window: no branch between the loads
      }
#pragma unroll
      for (uint32_t u = 0; u < kLoads; u++) {
        const uint32_t q = q0 + u * 64 + lane;
        if (q < words) sh[q] = t[u];
      }
    }
    __syncthreads();
    const uint32_t hi = min(e1, c0 + kReduceWindow);
    for (uint32_t e = max(e0, c0); e < hi; e++) sum += sh[(e - c0) * NS + k];
  }
  if (!live) return;
  // splat_proj_grad holds slots 0-8 in slot order
  if (k < 9)
    reinterpret_cast<float*>(grads + row)[k] = sum;
  else
    abs_grads[2 * (size_t)row + (k - 9)] = sum;
}

}  // namespace

namespace {

// Workgroups of a persistent launch: `per_cu` per compute unit.
unsigned persistent_grid(splat_ctx* ctx, unsigned per_cu, unsigned tiles) {
  int cus = 0;
  SPLAT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  return std::max(1u, std::min<unsigned>(tiles, (unsigned)cus * std::max(per_cu, 1u)));
}

template <int B, bool BATCH, bool PERSIST>
void fwd_go(splat_ctx* ctx, const ProjBuffers& pb, const BinBuffers& bb, unsigned T,
            uint32_t width, uint32_t height, const float bg[3], float* image, float* final_t,
            uint32_t* n_contrib, const BatchDev& bd, uint32_t* counter, unsigned per_cu) {
  if (PERSIST) fill_u32(ctx, counter, kXcds, 0u);
  const unsigned grid = PERSIST ? persistent_grid(ctx, per_cu, T) : T;
  launch(ctx, "raster_fwd", raster_fwd_kernel<B, BATCH, PERSIST>, dim3(grid), dim3(256), 0,
         (const float4*)pb.rec, (const uint32_t*)bb.tile_start, (const uint32_t*)bb.tile_order,
         (const uint32_t*)bb.sorted_gauss, T, bb.tiles_x, width, height, bg[0], bg[1], bg[2],
         image, final_t, n_contrib, bd, counter, (unsigned long long*)nullptr);
}

template <bool ORDERED, bool ABS, int BF, bool BATCH, bool PERSIST, bool ROWS>
void bwd_go(splat_ctx* ctx, const char* name, const ProjBuffers& pb, const BinBuffers& bb,
            unsigned T, uint32_t width, uint32_t height, const float bg[3], const float* final_t,
            const uint32_t* n_contrib, const float* dimage, float* slots, const BatchDev& bd,
            uint32_t* counter, unsigned per_cu) {
  if (PERSIST) fill_u32(ctx, counter, kXcds, 0u);
  const unsigned grid = PERSIST ? persistent_grid(ctx, per_cu, T) : T;
  launch(ctx, name, raster_bwd_kernel<ORDERED, ABS, BF, BATCH, PERSIST, ROWS>, dim3(grid), dim3(256),
         0, (const float4*)pb.rec, (const uint32_t*)bb.tile_start,
         (const uint32_t*)bb.tile_order, (const uint32_t*)bb.sorted_gauss,
         (const uint32_t*)bb.sorted_entry, T, bb.tiles_x, width,
         height, bg[0], bg[1], bg[2], final_t, n_contrib, dimage, slots, bd, counter,
         (unsigned long long*)nullptr);
}

template <bool BATCH, bool PERSIST>
void fwd_pick(splat_ctx* ctx, const RasterVariant& rv, const ProjBuffers& pb,
              const BinBuffers& bb, unsigned T, uint32_t width, uint32_t height,
              const float bg[3], float* image, float* final_t, uint32_t* n_contrib,
              const BatchDev& bd, uint32_t* counter) {
  switch (rv.fwd_batch) {
    case 32:
      return fwd_go<32, BATCH, PERSIST>(ctx, pb, bb, T, width, height, bg, image, final_t,
                                        n_contrib, bd, counter, rv.per_cu);
    case 128:
      return fwd_go<128, BATCH, PERSIST>(ctx, pb, bb, T, width, height, bg, image, final_t,
                                         n_contrib, bd, counter, rv.per_cu);
    case 256:  // one entry per lane: a whole batch per barrier pair
      return fwd_go<256, BATCH, PERSIST>(ctx, pb, bb, T, width, height, bg, image, final_t,
                                         n_contrib, bd, counter, rv.per_cu);
    default:
      return fwd_go<64, BATCH, PERSIST>(ctx, pb, bb, T, width, height, bg, image, final_t,
                                        n_contrib, bd, counter, rv.per_cu);
  }
}

template <bool BATCH, bool PERSIST, bool ROWS>
void bwd_rows_pick(splat_ctx* ctx, const RasterVariant& rv, const ProjBuffers& pb,
                   const BinBuffers& bb, unsigned T, uint32_t width, uint32_t height,
                   const float bg[3], const float* final_t, const uint32_t* n_contrib,
                   const float* dimage, float* slots, bool need_abs, const BatchDev& bd,
                   uint32_t* counter) {
  const unsigned pc = rv.per_cu;
  if (rv.bwd_batch == 16) {
    if (need_abs)
      return bwd_go<false, true, 16, BATCH, PERSIST, ROWS>(ctx, "raster_bwd", pb, bb, T, width,
                                                           height, bg, final_t, n_contrib,
                                                           dimage, slots, bd, counter, pc);
    return bwd_go<false, false, 16, BATCH, PERSIST, ROWS>(ctx, "raster_bwd", pb, bb, T, width,
                                                          height, bg, final_t, n_contrib, dimage,
                                                          slots, bd, counter, pc);
  }
  if (need_abs)
    return bwd_go<false, true, 32, BATCH, PERSIST, ROWS>(ctx, "raster_bwd", pb, bb, T, width,
                                                         height, bg, final_t, n_contrib, dimage,
                                                         slots, bd, counter, pc);
  return bwd_go<false, false, 32, BATCH, PERSIST, ROWS>(ctx, "raster_bwd", pb, bb, T, width,
                                                        height, bg, final_t, n_contrib, dimage,
                                                        slots, bd, counter, pc);
}

template <bool BATCH, bool PERSIST>
void bwd_pick(splat_ctx* ctx, const RasterVariant& rv, const ProjBuffers& pb,
              const BinBuffers& bb, unsigned T, uint32_t width, uint32_t height,
              const float bg[3], const float* final_t, const uint32_t* n_contrib,
              const float* dimage, bool ordered, float* slots, bool need_abs, const BatchDev& bd,
              uint32_t* counter) {
  const unsigned pc = rv.per_cu;
  // the record width (9 or 11 words per entry) follows need_abs in both modes
  if (ordered && need_abs)
    return bwd_go<true, true, 8, BATCH, PERSIST, false>(ctx, "raster_bwd_ordered", pb, bb, T,
                                                        width, height, bg, final_t, n_contrib,
                                                        dimage, slots, bd, counter, pc);
  if (ordered)
    return bwd_go<true, false, 8, BATCH, PERSIST, false>(ctx, "raster_bwd_ordered", pb, bb, T,
                                                         width, height, bg, final_t, n_contrib,
                                                         dimage, slots, bd, counter, pc);
  if (rv.rows)
    return bwd_rows_pick<BATCH, PERSIST, true>(ctx, rv, pb, bb, T, width, height, bg, final_t,
                                               n_contrib, dimage, slots, need_abs, bd, counter);
  return bwd_rows_pick<BATCH, PERSIST, false>(ctx, rv, pb, bb, T, width, height, bg, final_t,
                                              n_contrib, dimage, slots, need_abs, bd, counter);
}

}  // namespace

void raster_forward_counted(splat_ctx* ctx, const ProjBuffers& pb, const BinBuffers& bb,
                            uint32_t width, uint32_t height, const float bg[3], float* image,
                            float* final_t, uint32_t* n_contrib, unsigned long long* pairs) {
  const unsigned T = bb.tiles_x * bb.tiles_y;
  fill_u32(ctx, reinterpret_cast<uint32_t*>(pairs), 2 * kNumPairCounters, 0u);
  launch(ctx, "raster_fwd_count", raster_fwd_kernel<64, false, false, true>, dim3(T), dim3(256),
         0, (const float4*)pb.rec, (const uint32_t*)bb.tile_start, (const uint32_t*)nullptr,
         (const uint32_t*)bb.sorted_gauss, T, bb.tiles_x, width, height, bg[0], bg[1], bg[2],
         image, final_t, n_contrib, BatchDev{}, (uint32_t*)nullptr, pairs);
}

void raster_backward_counted(splat_ctx* ctx, const ProjBuffers& pb, const BinBuffers& bb,
                             uint32_t width, uint32_t height, const float bg[3],
                             const float* final_t, const uint32_t* n_contrib,
                             const float* dimage, float* slots, unsigned long long* pairs) {
  const unsigned T = bb.tiles_x * bb.tiles_y;
  launch(ctx, "raster_bwd_count", raster_bwd_kernel<false, false, 32, false, false, true, true>,
         dim3(T), dim3(256), 0, (const float4*)pb.rec, (const uint32_t*)bb.tile_start,
         (const uint32_t*)nullptr, (const uint32_t*)bb.sorted_gauss,
         (const uint32_t*)bb.sorted_entry, T, bb.tiles_x, width, height, bg[0], bg[1], bg[2],
         (const float*)final_t, (const uint32_t*)n_contrib, dimage, slots, BatchDev{},
         (uint32_t*)nullptr, pairs);
}

void raster_forward(splat_ctx* ctx, const ProjBuffers& pb, const BinBuffers& bb, uint32_t width,
                    uint32_t height, const float bg[3], float* image, float* final_t,
                    uint32_t* n_contrib, const RasterVariant& rv, const BatchDev* bd,
                    uint32_t* counter) {
  const unsigned T = bb.tiles_x * bb.tiles_y;
  const BatchDev none{};
  SPLAT_REQUIRE(!rv.persistent || counter, "persistent launch needs a counter");
  if (bd) {
    if (rv.persistent)
      fwd_pick<true, true>(ctx, rv, pb, bb, T, width, height, bg, image, final_t, n_contrib, *bd,
                           counter);
    else
      fwd_pick<true, false>(ctx, rv, pb, bb, T, width, height, bg, image, final_t, n_contrib,
                            *bd, counter);
  } else if (rv.persistent) {
    fwd_pick<false, true>(ctx, rv, pb, bb, T, width, height, bg, image, final_t, n_contrib, none,
                          counter);
  } else {
    fwd_pick<false, false>(ctx, rv, pb, bb, T, width, height, bg, image, final_t, n_contrib,
                           none, counter);
  }
}

void raster_backward(splat_ctx* ctx, const ProjBuffers& pb, const BinBuffers& bb, uint32_t width,
                     uint32_t height, const float bg[3], const float* final_t,
                     const uint32_t* n_contrib, const float* dimage, bool ordered, float* slots,
                     bool need_abs, const RasterVariant& rv, const BatchDev* bd,
                     uint32_t* counter) {
  const unsigned T = bb.tiles_x * bb.tiles_y;
  const BatchDev none{};
  SPLAT_REQUIRE(!rv.persistent || counter, "persistent launch needs a counter");
  if (bd) {
    if (rv.persistent)
      bwd_pick<true, true>(ctx, rv, pb, bb, T, width, height, bg, final_t, n_contrib, dimage,
                           ordered, slots, need_abs, *bd, counter);
    else
      bwd_pick<true, false>(ctx, rv, pb, bb, T, width, height, bg, final_t, n_contrib, dimage,
                            ordered, slots, need_abs, *bd, counter);
  } else if (rv.persistent) {
    bwd_pick<false, true>(ctx, rv, pb, bb, T, width, height, bg, final_t, n_contrib, dimage,
                          ordered, slots, need_abs, none, counter);
  } else {
    bwd_pick<false, false>(ctx, rv, pb, bb, T, width, height, bg, final_t, n_contrib, dimage,
                           ordered, slots, need_abs, none, counter);
  }
}

RasterTune tune_raster(splat_ctx* ctx, const ProjBuffers& pb, const BinBuffers& bb,
                       uint32_t width, uint32_t height, const float bg[3], float* image,
                       float* final_t, uint32_t* n_contrib, const float* dimage, bool ordered,
                       float* slots, bool need_abs, const BatchDev* bd, uint32_t* counter) {
  hipEvent_t e0, e1;
  SPLAT_HIP(hipEventCreate(&e0));
  SPLAT_HIP(hipEventCreate(&e1));
  const bool timing = ctx->timing;
  ctx->timing = false;
  RasterTune out;
  try {
    auto time_it = [&](auto&& fn) {
      float best = 1e30f;
      for (int r = 0; r < 3; r++) {
        SPLAT_HIP(hipEventRecord(e0, ctx->stream));
        fn();
        SPLAT_HIP(hipEventRecord(e1, ctx->stream));
        SPLAT_HIP(hipEventSynchronize(e1));
        float ms = 0;
        SPLAT_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
      }
      return best;
    };
    const RasterVariant def;
    out.fwd_ms = out.bwd_ms = 1e30f;
    for (int fb : {32, 64, 128, 256})
      for (int mode = 0; mode < 4; mode++) {
        RasterVariant rv;
        rv.fwd_batch = fb;
        rv.persistent = mode > 0;
        rv.per_cu = mode == 1 ? 2 : (mode == 2 ? 4 : 8);
        const float ms = time_it([&] {
          raster_forward(ctx, pb, bb, width, height, bg, image, final_t, n_contrib, rv, bd,
                         counter);
        });
        out.candidates++;
        if (rv.fwd_batch == def.fwd_batch && !rv.persistent) out.default_fwd_ms = ms;
        if (ms < out.fwd_ms) {
          out.fwd_ms = ms;
          out.fwd = rv;
        }
      }
    for (int bw : {16, 32})
      for (int mode = 0; mode < 8; mode++) {
        RasterVariant rv;
        rv.bwd_batch = bw;
        rv.rows = mode < 4 || ordered;  // the ordered form walks per entry
        if (ordered && mode >= 4) continue;
        rv.persistent = mode % 4 > 0;
        rv.per_cu = mode % 4 == 1 ? 2 : (mode % 4 == 2 ? 4 : 8);
        const float ms = time_it([&] {
          raster_backward(ctx, pb, bb, width, height, bg, final_t, n_contrib, dimage, ordered,
                          slots, need_abs, rv, bd, counter);
        });
        out.candidates++;
        if (rv.bwd_batch == def.bwd_batch && !rv.persistent && rv.rows) out.default_bwd_ms = ms;
        if (ms < out.bwd_ms) {
          out.bwd_ms = ms;
          out.bwd = rv;
        }
      }
  } catch (...) {
    ctx->timing = timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  ctx->timing = timing;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return out;
}

void fill_tune_result(const RasterTune& t, splat_tune_result* out) {
  out->fwd_batch = t.fwd.fwd_batch;
  out->fwd_persistent = t.fwd.persistent;
  out->fwd_per_cu = (int32_t)t.fwd.per_cu;
  out->bwd_batch = t.bwd.bwd_batch;
  out->bwd_persistent = t.bwd.persistent;
  out->bwd_per_cu = (int32_t)t.bwd.per_cu;
  out->fwd_ms = t.fwd_ms;
  out->bwd_ms = t.bwd_ms;
  out->default_fwd_ms = t.default_fwd_ms;
  out->default_bwd_ms = t.default_bwd_ms;
  out->candidates = t.candidates;
  out->bwd_row_lists = t.bwd.rows ? 1 : 0;
}

void reduce_slots(splat_ctx* ctx, const uint32_t* entry_offset, const float* slots, size_t n,
                  splat_proj_grad* grads, float* abs_grads, const StepStatus* status) {
  static_assert(sizeof(splat_proj_grad) == 9 * sizeof(float), "slot order of splat_proj_grad");
  if (abs_grads)
    launch(ctx, "reduce_slots", reduce_slots_kernel<11>, dim3(blocks_for(n, 64 / 11)), dim3(64),
           0, entry_offset, slots, (uint32_t)n, grads, abs_grads, status);
  else
    launch(ctx, "reduce_slots", reduce_slots_kernel<9>, dim3(blocks_for(n, 64 / 9)), dim3(64), 0,
           entry_offset, slots, (uint32_t)n, grads, abs_grads, status);
}

}  // namespace splat
