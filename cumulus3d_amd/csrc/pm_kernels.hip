// pm_kernels.hip — gfx950 PatchStereo kernels (random init + evaluation, red-black half step,
// read-out filter, upsampling, median filter, cost probe) and their launchers.
//
// The kernels follow the rules of crates/dense/src/kernel.rs:36-74 as implemented by the reference's
// crates/cuda/src/kernels/patchmatch.cu (cited per kernel). Work decomposition (MI355X-specific):
//   * wave64 = 64 active pixels of one row (red-black: every other pixel), 1, 2 or 4 rows per
//     workgroup (autotuned); a workgroup stages its reference patch rows (128 + 2R px wide) in LDS once;
//   * all views of a batch share one flattened 1-D grid of tiles (no empty blocks for smaller
//     views), remapped so that consecutive tiles run on the same XCD and share its L2;
//   * the per-thread view weights of the refinement live in LDS, everything else in VGPRs; the
//     hypothesis and source loops stay rolled so that one inlined NCC body serves all hypotheses.
#include "pm_device.hpp"
#include "pm_kernels.hpp"

namespace c3d {

// rows of active pixels per half-step workgroup: a compile-time variant (1, 2 or 4), picked per
// session by the tile autotuner (c3d_pm_autotune). One row (a one-wave workgroup) is the default: a
// wave's slot on its SIMD is refilled as soon as that wave ends, where a 4-wave workgroup holds its
// slots (and its LDS) until its slowest row is done, and the rows' work differs through the
// data-dependent early exits.
constexpr int HALF_SPAN = 128;   // pixel span of one row of 64 active pixels
constexpr int PIX_TW = 64;       // eval/filter tiles: 64 x 4 pixels
constexpr int PIX_TH = 4;

// The run-start evaluation and the read-out keep five source gathers in flight through a window
// and, for register-held windows of at most 36 samples, ask for three waves per SIMD: 156-158
// VGPRs without spills (eight in flight needed 180-187 VGPRs, two waves per SIMD).
constexpr int EVAL_PIPE = 5;
template <int WR, int WSTEP>
constexpr int eval_waves() {
    return (WR > 0 && Win<(WR > 0 ? WR : 1), WSTEP>::WN2 <= 36) ? 3 : 1;
}

// region offsets for the up direction (patchmatch.cu:424-436)
__constant__ int V_OFF[7][2] = {{0, -1}, {-1, -2}, {1, -2}, {-2, -3}, {2, -3}, {-3, -4}, {3, -4}};

__device__ __forceinline__ void rot_dir(int dir, int dx, int dy, int& ox, int& oy) {
    ox = dx;
    oy = dy;
    for (int k = 0; k < dir; ++k) {
        int t = ox;
        ox = -oy;
        oy = t;
    }
}

template <int N>
__device__ __forceinline__ float4 pick4(const float4 (&a)[N], int i) {
    float4 r = a[0];
#pragma unroll
    for (int q = 1; q < N; ++q)
        if (q == i) r = a[q];
    return r;
}

// Reference tile accessor: compile-time pitch for the fixed windows, runtime pitch for WR = 0.
template <int WR, int TW>
using TileRefT = typename std::conditional<(WR > 0), RefTile<TW>, RefTileDyn>::type;

template <int WR, int TW>
__device__ __forceinline__ TileRefT<WR, TW> tile_ref(const float* tile, int x0, int y0, int pitch) {
    if constexpr (WR > 0) return RefTile<TW>{tile, x0, y0};
    else return RefTileDyn{tile, x0, y0, pitch};
}

// (photometric cost, truncated geometric error) of one hypothesis x one source with the pixel's
// reciprocal focal lengths hoisted (patchmatch.cu:316-324)
template <int WR, int WSTEP, bool GEO, int PIPE, bool AHEAD, class REF>
__device__ __forceinline__ float pair_cost_h(const KParams& P, const RefWinT<WR, WSTEP>& rw, const REF& ref, const LevelDesc& R,
                                             const LevelDesc& S, const PairRec& pr, const LevelPair& lp, float ifx, float ify, int x,
                                             int y, float4 h, float& geo) {
    float H[9];
    float c = 2.0f;
    // geometric runs: the snapshot depth is fetched before the NCC window and used after it
    GeoFetch gf{0.f, 0.f, 0.f, false};
    if constexpr (GEO) gf = geo_fetch(P, R, S, pr, x, y, h.w);
    if (homography(R, lp, ifx, ify, (float)x, (float)y, h.w, xyz(h), H)) c = ncc_cost_any<WR, WSTEP, PIPE, GEO, AHEAD>(P, rw, ref, S, H, x, y);
    geo = GEO ? fminf(geo_finish(R, S, pr, x, y, gf), P.gmax) : 0.f;
    return c;
}

// ================================================================ half step (patchmatch.cu:439-657)
// GEO: geometric-consistency run (the geometric term is compiled in only there).
// TPP: lanes per pixel. With 2, the two lanes of a pixel (adjacent lanes) keep their own copies of
// the pixel's state and split its NCC windows: the step-2 hypotheses by parity and the refinement
// candidates by parity, exchanging the costs with a lane swap; every decision is then taken on both
// lanes from the same values in the same order, so the results are those of one lane per pixel.
template <int WR, int WSTEP, bool GEO, int ROWS, int TPP>
__global__ __launch_bounds__(ROWS * 64) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_half(KParams P) {
    constexpr int NT = ROWS * 64;
    constexpr int SPAN = HALF_SPAN / TPP;  // pixel span of one wave row (64 / TPP active pixels)
    // source gathers in flight in the window loop: eight where the registers allow (photometric,
    // one lane per pixel: 247 VGPRs), seven in the geometric build (255), six with two lanes per
    // pixel, fewer for the 49 register-held weights of a 7 x 7 window (deeper pipelines spill)
    constexpr int PIPE = Win<(WR > 0 ? WR : 1), WSTEP>::WN2 > 36 ? 3 : (TPP > 1 ? 6 : (GEO ? 7 : 8));
    // the reference value of a window sample is read with its gather (see window_sums)
    constexpr bool AHEAD = Win<(WR > 0 ? WR : 1), WSTEP>::WN2 <= 36;
    constexpr int TW = SPAN + 2 * WR;
    constexpr int TH = ROWS + 2 * WR;
    __shared__ float tile_s[WR > 0 ? TW * TH : 1];
    // per-thread hypothesis table: the 9 candidates of step 2, then the refinement candidates
    __shared__ float4 hyp_lds[9 * NT];
    // dynamic LDS: [reference tile (WR = 0 only)] [per-thread view weights]
    extern __shared__ float dyn_lds[];
    const int wr = WR > 0 ? WR : P.wr;
    const int twd = SPAN + 2 * wr, thd = ROWS + 2 * wr;
    float* tile = WR > 0 ? tile_s : dyn_lds;
    float* wv_lds = WR > 0 ? dyn_lds : dyn_lds + twd * thd;

    BatchItem it;
    int px0, py0;
    if (!find_tile(P, xcd_tile(blockIdx.x, P.ntiles), SPAN, ROWS, it, px0, py0)) return;
    const LevelDesc R = P.lv[it.view * P.nlevels + P.level];
    const int w = it.w, h = it.h;
    {
        const uint8_t* g = P.gray + R.goff;
        const int tx0 = px0 - wr, ty0 = py0 - wr;
        for (int i = threadIdx.x; i < twd * thd; i += NT) {
            int ty = i / twd, tx = i - ty * twd;
            int gx = tx0 + tx, gy = ty0 + ty;
            tile[i] = (gx >= 0 && gy >= 0 && gx < w && gy < h) ? (float)g[(size_t)gy * w + gx] : 0.0f;
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const int sub = lane & (TPP - 1);  // this lane's share of the pixel's windows
    const int y = py0 + row;
    const int x = px0 + 2 * (lane / TPP) + ((y + P.color) & 1);
    if (x >= w || y >= h) return;  // both lanes of a pixel leave together

    const auto ref = tile_ref<WR, TW>(tile, px0 - wr, py0 - wr, twd);
    const ViewDesc vd = P.views[it.view];
    const int nsrc = vd.nsrc;
    const PairRec* prs = P.pairs + (size_t)it.view * MAXSRC;
    const LevelPair* lps = P.lpairs + ((size_t)it.view * P.nlevels + P.level) * MAXSRC;
    const long long base = it.off;
    const int pix = y * w + x;
    float4* pl = P.plane + base;
    const float* co = P.cost + base;
    const f3 ray = pixel_ray(R, x, y);
    const float ifx = 1.0f / R.fx, ify = 1.0f / R.fy;
    float* wv = wv_lds + threadIdx.x;
    float4* hyp = hyp_lds + threadIdx.x;  // hyp[i * NT]

    // 1. candidates: current + best of 8 regions (patchmatch.cu:477-511). The 72 neighbour costs
    //    are loaded unconditionally (out-of-image offsets read the pixel's own cost and are ignored).
    const float4 hyp0 = pl[pix];
    hyp[0] = hyp0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int dir = r & 3;
        const bool strip = r >= 4;
        const int cnt = strip ? 11 : 7;
        float best = 1e30f;
        int bq = -1, bx = 0, by = 0;
#pragma unroll
        for (int k = 0; k < 11; ++k) {
            if (k >= cnt) break;
            int dx, dy;
            if (strip) rot_dir(dir, 0, -(3 + 2 * k), dx, dy);
            else rot_dir(dir, V_OFF[k][0], V_OFF[k][1], dx, dy);
            const int qx = x + dx, qy = y + dy;
            const bool inb = qx >= 0 && qy >= 0 && qx < w && qy < h;
            const int qi = inb ? qy * w + qx : pix;
            const float cq = co[qi];
            if (inb && cq < best) {
                best = cq;
                bq = qi;
                bx = qx;
                by = qy;
            }
        }
        float4 hr = hyp0;
        if (bq >= 0) {
            float4 q = pl[bq];
            f3 n = xyz(q);
            f3 rq = pixel_ray(R, bx, by);
            float den = dot3(n, ray);
            if (den < -1e-6f && q.w > 0.f) {
                float d = q.w * dot3(n, rq) / den;
                if (d > 0.f && isfinite(d)) hr = mk4(n, d);
            }
        }
        hyp[(1 + r) * NT] = hr;
    }

    // 2. view selection and aggregation (patchmatch.cu:513-594)
    RefWinT<WR, WSTEP> rw;
    ref_stats_any<WR, WSTEP>(P, ref, x, y, wr, rw);
    const int vprev = P.bview[base + pix];
    float num[9], numf[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) num[i] = numf[i] = 0.f;
    float den = 0.f, wbest = 0.f, denf = 0.f;
    int vbest = 255;
#pragma nounroll
    for (int j = 0; j < nsrc; ++j) {
        const PairRec& pr = prs[j];
        const LevelDesc S = P.lv[pr.src * P.nlevels + P.level];
        const LevelPair& lp = lps[j];
        float prior;
        {
            float H[9];
            bool ok = homography(R, lp, ifx, ify, (float)x, (float)y, hyp0.w, xyz(hyp0), H);
            prior = view_prior(P, R, pr, H, ok, x, y, hyp0, (float)wr);
        }
        wv[j * P.wv_stride] = 0.f;
        if (!(prior > 0.f)) continue;
        float m[9], gg[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = gg[q] = 0.f;
#pragma nounroll
        for (int i = sub; i < 9; i += TPP) {
            float gi;
            const float mi = pair_cost_h<WR, WSTEP, GEO, PIPE, AHEAD>(P, rw, ref, R, S, pr, lp, ifx, ify, x, y, hyp[i * NT], gi);
#pragma unroll
            for (int q = 0; q < 9; ++q)
                if (q == i) {
                    m[q] = mi;
                    gg[q] = gi;
                }
        }
        if constexpr (TPP > 1) {
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const float om = __shfl_xor(m[q], 1), og = __shfl_xor(gg[q], 1);
                if ((q & 1) != sub) {
                    m[q] = om;
                    gg[q] = og;
                }
            }
        }
        int good = 0, bad = 0;
        float conf = 0.f;
#pragma unroll
        for (int i = 1; i < 9; ++i) {
            if (m[i] < P.tau_good) {
                good++;
                conf += c3d_expf(-m[i] * m[i] * P.inv2beta2);
            }
            if (m[i] > P.tau_bad) bad++;
        }
        bool adopt = good > P.n1 && bad < P.n2;
        bool isprev = j == vprev;
        float wp = adopt ? (isprev ? 2.f : 1.f) * conf / (float)good : (isprev ? P.prev_w : 0.f);
        float w2 = wp * prior;
        wv[j * P.wv_stride] = w2 > 0.f ? w2 : -prior;
#pragma unroll
        for (int i = 0; i < 9; ++i) numf[i] += prior * (m[i] + P.lambda * gg[i]);
        denf += prior;
        if (w2 > 0.f) {
#pragma unroll
            for (int i = 0; i < 9; ++i) num[i] += w2 * (m[i] + P.lambda * gg[i]);
            den += w2;
            if (w2 > wbest) {
                wbest = w2;
                vbest = j;
            }
        }
    }
    float agg[9];
    int bi9 = 0;
    if (den > 0.f) {
#pragma unroll
        for (int i = 0; i < 9; ++i) agg[i] = num[i] / den;
#pragma unroll
        for (int i = 1; i < 9; ++i)
            if (agg[i] < agg[bi9]) bi9 = i;
    } else if (denf > 0.f) {
#pragma unroll
        for (int i = 0; i < 9; ++i) agg[i] = numf[i] / denf;
#pragma unroll
        for (int i = 1; i < 9; ++i)
            if (agg[i] < agg[bi9]) bi9 = i;
        for (int j = 0; j < nsrc; ++j) wv[j * P.wv_stride] = -wv[j * P.wv_stride];
        den = denf;
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) agg[i] = 2.f;
    }
    float4 best = hyp[bi9 * NT];
    float bcost = 2.f;
#pragma unroll
    for (int q = 0; q < 9; ++q)
        if (q == bi9) bcost = agg[q];

    // 3. refinement: 6 hypotheses, plus the coarse prior at weak texture (patchmatch.cu:596-653)
    unsigned long long st = rng_state(P.seed, vd.key, P.level, P.run_id, P.t, P.color, (unsigned long long)pix);
    float d_r = random_depth(st, vd.dmin, vd.dmax);
    f3 n_r = random_normal(st, ray);
    float u = 2.f * rnd(st) - 1.f;
    float inv = best.w > 0.f ? (1.f / best.w) * (1.f + u * P.eps) : -1.f;
    float d_p = inv > 0.f ? 1.f / inv : d_r;
    if (!(best.w > 0.f)) best.w = d_r;
    f3 n0 = xyz(best);
    f3 n_p = perturb_normal(st, n0, ray, P.phi);
    float4* cand = hyp;  // the hypothesis table is free again
    cand[0] = mk4(n0, d_r);
    cand[NT] = mk4(n_r, best.w);
    cand[2 * NT] = mk4(n_r, d_r);
    cand[3 * NT] = mk4(n0, d_p);
    cand[4 * NT] = mk4(n_p, best.w);
    cand[5 * NT] = mk4(n_p, d_p);
    int ncand = 6;
    if (P.use_prior && rw.var < P.weak_var255) {
        float4 q = P.prior[base + pix];
        if (q.w > 0.f && dot3(xyz(q), ray) < 0.f) {
            cand[6 * NT] = q;
            ncand = 7;
        }
    }
    // A candidate's cost is a sum of non-negative terms (weights > 0, costs in [0, 2], geometric
    // errors in [0, gmax], lambda >= 0) divided by den: its partial sums only grow, so once
    // nm / den reaches bcost the candidate cannot be accepted and its remaining sources are
    // skipped. The accepted planes and costs are unchanged.
    // With two lanes per pixel each lane costs every other candidate against the step-2 best (a
    // candidate that reaches it cannot be accepted later either, the best only decreasing), then
    // both lanes accept in candidate order from the exchanged costs.
    const bool early = P.lambda >= 0.f && (!GEO || P.gmax >= 0.f);
    float ccs[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) ccs[q] = 3.f;  // above any cost: not accepted
#pragma nounroll
    for (int c = sub; c < ncand; c += TPP) {
        const float4 hc = cand[c * NT];
        float cc = 2.f;
        bool out = false;
        if (den > 0.f) {
            float nm = 0.f;
#pragma nounroll
            for (int j = 0; j < nsrc; ++j) {
                float wj = wv[j * P.wv_stride];
                if (!(wj > 0.f)) continue;
                const PairRec& pr = prs[j];
                const LevelDesc S = P.lv[pr.src * P.nlevels + P.level];
                float g;
                float m = pair_cost_h<WR, WSTEP, GEO, PIPE, AHEAD>(P, rw, ref, R, S, pr, lps[j], ifx, ify, x, y, hc, g);
                nm += wj * (m + P.lambda * g);
                if (early && nm / den >= bcost) {
                    out = true;
                    break;
                }
            }
            cc = nm / den;
        }
        if constexpr (TPP == 1) {
            if (cc < bcost) {
                bcost = cc;
                best = hc;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 7; ++q)
                if (q == c) ccs[q] = out ? 3.f : cc;
        }
    }
    if constexpr (TPP > 1) {
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            const float o = __shfl_xor(ccs[q], 1);
            if ((q & 1) != sub) ccs[q] = o;
        }
#pragma unroll
        for (int q = 0; q < 7; ++q)
            if (q < ncand && ccs[q] < bcost) {
                bcost = ccs[q];
                best = cand[q * NT];
            }
    }
    if (sub == 0) {
        pl[pix] = best;
        P.cost[base + pix] = bcost;
        P.bview[base + pix] = (uint8_t)vbest;
    }
}

// Stages the reference rows of a 64 x 4 pixel tile (plus the window margin) in LDS.
__device__ __forceinline__ void stage_pix_tile(float* tile, const uint8_t* g, int w, int h, int px0, int py0, int wr) {
    const int TW = PIX_TW + 2 * wr, TH = PIX_TH + 2 * wr;
    const int tx0 = px0 - wr, ty0 = py0 - wr;
    for (int i = threadIdx.x; i < TW * TH; i += 256) {
        int ty = i / TW, tx = i - ty * TW;
        int gx = tx0 + tx, gy = ty0 + ty;
        tile[i] = (gx >= 0 && gy >= 0 && gx < w && gy < h) ? (float)g[(size_t)gy * w + gx] : 0.0f;
    }
}

// ================================================================ run start / evaluation (patchmatch.cu:659-692)
//
// A run starts with pm_eval: random initialisation (mode 1, by random_init), the top-K cost of
// every pixel's plane and the best view reset. When the run then has at least one iteration, the
// costs of the pixels of colour 0 are dead: the first half step (colour 0) recomputes and rewrites
// every colour-0 pixel without reading its stored cost (its candidate step reads only the 72
// neighbours of the opposite colour), and nothing else reads them in between. So such a run starts
// with k_eval_init over every pixel (the initialisation and best-view reset, no window) and
// k_eval<ONE = true> for the pixels of colour 1 only, with the half step's one-colour lane
// mapping, which does half the window work. Runs of no iterations and c3d_pm_evaluate keep the
// whole evaluation (ONE = false).
__device__ __forceinline__ float4 eval_init_plane(const KParams& P, const ViewDesc& vd, const LevelDesc& R, int x, int y, int pix,
                                                  long long base, int mode) {
    float4 h = P.plane[base + pix];
    // random_init 1: every pixel (patchmatch.cu:668-676); 2: pixels without a positive depth
    if (mode == 1 && (P.random_init == 1 || (P.random_init == 2 && !(h.w > 0.f)))) {
        unsigned long long st = rng_state(P.seed, vd.key, P.level, P.run_id, 0xFFFF, 0, (unsigned long long)pix);
        float d = random_depth(st, vd.dmin, vd.dmax);
        f3 n = random_normal(st, pixel_ray(R, x, y));
        h = mk4(n, d);
        P.plane[base + pix] = h;
    }
    return h;
}

// top-K mean of the plane's per-source costs (patchmatch.cu:326-345), 2 for a non-positive depth
template <int WR, int WSTEP, class REF>
__device__ __forceinline__ float eval_cost(const KParams& P, const ViewDesc& vd, const LevelDesc& R, const PairRec* prs,
                                           const LevelPair* lps, const REF& ref, int wr, int x, int y, float4 h) {
    if (!(h.w > 0.f)) return 2.f;
    RefWinT<WR, WSTEP> rw;
    ref_stats_any<WR, WSTEP>(P, ref, x, y, wr, rw);
    float b0 = 1e30f, b1 = 1e30f, b2 = 1e30f;
#pragma nounroll
    for (int j = 0; j < vd.nsrc; ++j) {
        const PairRec& pr = prs[j];
        const LevelDesc S = P.lv[pr.src * P.nlevels + P.level];
        float gq = 0.f;
        float cc = pair_cost<WR, WSTEP, EVAL_PIPE>(P, rw, ref, R, S, pr, lps[j], x, y, h, &gq);
        if (P.geometric) cc = cc + P.lambda * gq;
        if (cc < b0) {
            b2 = b1;
            b1 = b0;
            b0 = cc;
        } else if (cc < b1) {
            b2 = b1;
            b1 = cc;
        } else if (cc < b2) {
            b2 = cc;
        }
    }
    int k = P.topk < vd.nsrc ? P.topk : vd.nsrc;
    float s = b0;
    if (k >= 2) s += b1;
    if (k >= 3) s += b2;
    return k > 0 ? s / (float)k : 2.0f;
}

// ONE = false: every pixel of 64 x 4 tiles, initialisation (mode 1) and cost. ONE = true: the
// pixels of colour P.color of 128 x 4 spans (one lane per pixel, as the half step), cost only
// (k_eval_init has initialised them).
template <int WR, int WSTEP, bool ONE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(eval_waves<WR, WSTEP>()))) void k_eval(KParams P, int mode) {
    constexpr int SPAN = ONE ? HALF_SPAN : PIX_TW;
    constexpr int TW = SPAN + 2 * WR;
    __shared__ float tile_s[WR > 0 ? TW * (PIX_TH + 2 * WR) : 1];
    extern __shared__ float dyn_lds[];
    const int wr = WR > 0 ? WR : P.wr;
    float* tile = WR > 0 ? tile_s : dyn_lds;
    BatchItem it;
    int px0, py0;
    if (!find_tile(P, xcd_tile(blockIdx.x, P.ntiles), SPAN, PIX_TH, it, px0, py0)) return;
    const LevelDesc R = P.lv[it.view * P.nlevels + P.level];
    {
        const uint8_t* g = P.gray + R.goff;
        const int twd = SPAN + 2 * wr, thd = PIX_TH + 2 * wr;
        const int tx0 = px0 - wr, ty0 = py0 - wr;
        for (int i = threadIdx.x; i < twd * thd; i += 256) {
            int ty = i / twd, tx = i - ty * twd;
            int gx = tx0 + tx, gy = ty0 + ty;
            tile[i] = (gx >= 0 && gy >= 0 && gx < it.w && gy < it.h) ? (float)g[(size_t)gy * it.w + gx] : 0.0f;
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, y = py0 + (threadIdx.x >> 6);
    const int x = ONE ? px0 + 2 * lane + ((y + P.color) & 1) : px0 + lane;
    if (x >= it.w || y >= it.h) return;
    const ViewDesc vd = P.views[it.view];
    const PairRec* prs = P.pairs + (size_t)it.view * MAXSRC;
    const LevelPair* lps = P.lpairs + ((size_t)it.view * P.nlevels + P.level) * MAXSRC;
    const int pix = y * it.w + x;
    const long long base = it.off;
    const float4 h = ONE ? P.plane[base + pix] : eval_init_plane(P, vd, R, x, y, pix, base, mode);
    const auto ref = tile_ref<WR, TW>(tile, px0 - wr, py0 - wr, SPAN + 2 * wr);
    P.cost[base + pix] = eval_cost<WR, WSTEP>(P, vd, R, prs, lps, ref, wr, x, y, h);
    if (!ONE && mode == 1) P.bview[base + pix] = 255;
}

// The batch item holding pixel g of the flattened batch, by a binary search over the items' pixel
// offsets (stored in increasing order); false for a pixel outside every item. Per-pixel passes use
// it instead of locate_tile: a few loads per thread, where a workgroup's linear search over the
// items for its tile costs more than a light pass itself on a 108-view batch.
__device__ __forceinline__ bool locate_pixel(const KParams& P, long long g, BatchItem& it, int& item, int& pix) {
    int lo = 0, hi = P.nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (P.batch[mid].off <= g) lo = mid;
        else hi = mid - 1;
    }
    it = P.batch[lo];
    item = lo;
    if (g < it.off || g - it.off >= (long long)it.w * it.h) return false;
    pix = (int)(g - it.off);
    return true;
}

// The window-free part of a run start for every pixel: random initialisation by random_init and
// the best-view reset (patchmatch.cu:668-677). Stored costs are not touched. One thread per pixel
// of the flattened batch.
__global__ __launch_bounds__(256) void k_eval_init(KParams P, long long npix) {
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= npix) return;
    BatchItem it;
    int item, pix;
    if (!locate_pixel(P, g, it, item, pix)) return;
    const int y = pix / it.w, x = pix - y * it.w;
    const LevelDesc R = P.lv[it.view * P.nlevels + P.level];
    const ViewDesc vd = P.views[it.view];
    eval_init_plane(P, vd, R, x, y, pix, it.off, 1);
    P.bview[it.off + pix] = 255;
}

// ================================================================ final read-out (patchmatch.cu:694-737)
template <int WR, int WSTEP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(eval_waves<WR, WSTEP>()))) void k_filter(KParams P) {
    constexpr int TW = PIX_TW + 2 * WR;
    __shared__ float tile_s[WR > 0 ? TW * (PIX_TH + 2 * WR) : 1];
    extern __shared__ float dyn_lds[];
    const int wr = WR > 0 ? WR : P.wr;
    float* tile = WR > 0 ? tile_s : dyn_lds;
    BatchItem it;
    int px0, py0;
    if (!find_tile(P, xcd_tile(blockIdx.x, P.ntiles), PIX_TW, PIX_TH, it, px0, py0)) return;
    const LevelDesc R = P.lv[it.view * P.nlevels + P.level];
    stage_pix_tile(tile, P.gray + R.goff, it.w, it.h, px0, py0, wr);
    __syncthreads();
    const int x = px0 + (threadIdx.x & 63), y = py0 + (threadIdx.x >> 6);
    if (x >= it.w || y >= it.h) return;
    const ViewDesc vd = P.views[it.view];
    const PairRec* prs = P.pairs + (size_t)it.view * MAXSRC;
    const LevelPair* lps = P.lpairs + ((size_t)it.view * P.nlevels + P.level) * MAXSRC;
    const int pix = y * it.w + x;
    const long long base = it.off;
    float4 h = P.plane[base + pix];
    int count = 0;
    if (h.w > 0.f) {
        const auto ref = tile_ref<WR, TW>(tile, px0 - wr, py0 - wr, PIX_TW + 2 * wr);
        RefWinT<WR, WSTEP> rw;
        ref_stats_any<WR, WSTEP>(P, ref, x, y, wr, rw);
        f3 ray = pixel_ray(R, x, y);
        f3 X = mk3(ray.x * h.w, ray.y * h.w, h.w);
        float lx = sqrtf(dot3(X, X));
#pragma nounroll
        for (int j = 0; j < vd.nsrc; ++j) {
            const PairRec& pr = prs[j];
            const LevelDesc S = P.lv[pr.src * P.nlevels + P.level];
            f3 Sv = mk3(pr.c[0] - X.x, pr.c[1] - X.y, pr.c[2] - X.z);
            float ls = sqrtf(dot3(Sv, Sv));
            if (!(ls > 0.f)) continue;
            float cost = -dot3(Sv, X) / (ls * lx);
            if (cost > P.cos_filter_tri) continue;
            float cphi = dot3(Sv, xyz(h)) / ls;
            if (cphi <= 0.f) continue;
            float H[9];
            if (!homography(R, lps[j], 1.0f / R.fx, 1.0f / R.fy, (float)x, (float)y, h.w, xyz(h), H)) continue;
            const GeoFetch gf = geo_fetch(P, R, S, pr, x, y, h.w);  // used after the window
            float m = ncc_cost_any<WR, WSTEP, EVAL_PIPE>(P, rw, ref, S, H, x, y);
            float e = P.ncc_norm * c3d_expf(-m * m * P.inv2ncc);
            if (e / (e + 0.5f) < P.q_min) continue;
            if (geo_finish(R, S, pr, x, y, gf) > P.filter_gmax) continue;
            count++;
        }
    }
    P.count[base + pix] = (uint8_t)(count < 255 ? count : 255);
}

// ================================================================ cost probe (no reference counterpart)
template <int WR, int WSTEP>
__global__ __launch_bounds__(256) void k_pair_costs(KParams P, int view, int slot, int n, const int* xy, const float4* planes,
                                                    float* out) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const LevelDesc R = P.lv[view * P.nlevels + P.level];
    const PairRec& pr = P.pairs[(size_t)view * MAXSRC + slot];
    const LevelPair& lp = P.lpairs[((size_t)view * P.nlevels + P.level) * MAXSRC + slot];
    const LevelDesc S = P.lv[pr.src * P.nlevels + P.level];
    int x = xy[2 * i], y = xy[2 * i + 1];
    RefGlobal g{P.gray + R.goff, R.w, R.h};
    const int wr = WR > 0 ? WR : P.wr;
    RefWinT<WR, WSTEP> rw;
    ref_stats_any<WR, WSTEP>(P, g, x, y, wr, rw);
    float4 h = planes[i];
    float H[9];
    float c = 2.f;
    if (homography(R, lp, 1.0f / R.fx, 1.0f / R.fy, (float)x, (float)y, h.w, xyz(h), H)) c = ncc_cost_any<WR, WSTEP>(P, rw, g, S, H, x, y);
    out[i] = c;
}

// ================================================================ plane records in LDS
//
// Both scale-change kernels below transfer a neighbour's plane onto a pixel's own viewing ray.
// The neighbour q (plane normal n_q, depth d_q on its ray r_q) holds the plane n_q . X = c_q with
// c_q = d_q (n_q . r_q), so the depth it gives pixel p is c_q / (n_q . r_p) (patchmatch.cu:405-412
// transfer, same f32 operations and order). A workgroup therefore stages, once, the records
// (n_q, c_q) of every neighbour its pixels can reach; c_q is NaN for a missing neighbour (outside
// the image, or d_q not > 0), which every later test rejects exactly as the explicit checks would.

__device__ __forceinline__ float4 plane_record(float4 q, f3 ray_q) {
    const f3 n = xyz(q);
    const float c = q.w > 0.f ? q.w * dot3(n, ray_q) : __int_as_float(0x7fc00000);
    return mk4(n, c);
}

// ================================================================ joint bilateral upsampling
// Output rules of pm_upsample (patchmatch.cu:739-801, upsample.rs:104-175): every low-resolution
// neighbour in the 5 x 5 block around the pixel's position at the lower scale whose transferred depth
// is positive and finite contributes exp(-ds/(2 sigma_s^2) - (I_p - I_q)^2/(2 sigma_c^2)) in scan
// order; with no weight the nearest such neighbour (first in scan order on ties) is taken; the
// normal is normalised and turned towards the camera.
//
// A 64 x 4 pixel tile reaches the low-resolution columns [x0/2 - 2, x0/2 + 33] and rows
// [y0/2 - 2, y0/2 + 3]: 36 x 6 plane records and gray levels are staged in LDS by the whole
// workgroup, so each of the 25 taps of a pixel is two LDS reads instead of a float4 and a byte
// gather from global memory with its own transfer arithmetic.
constexpr int UP_CW = PIX_TW / 2 + 4;  // 36 staged low-resolution columns
constexpr int UP_CH = PIX_TH / 2 + 4;  // 6 staged rows

__global__ __launch_bounds__(256) void k_upsample(KParams P) {
    __shared__ float4 rec[UP_CH * UP_CW];
    __shared__ float lgray[UP_CH * UP_CW];
    BatchItem it;
    int px0, py0;
    if (!find_tile(P, xcd_tile(blockIdx.x, P.ntiles), PIX_TW, PIX_TH, it, px0, py0)) return;  // uniform per workgroup
    const LevelDesc R = P.lv[it.view * P.nlevels + P.level];
    const LevelDesc L = P.lv[it.view * P.nlevels + P.level - 1];
    const float4* low = P.prior + it.low_off;
    const int qxa = px0 / 2 - 2, qya = py0 / 2 - 2;  // first staged low-resolution column and row
    for (int k = threadIdx.x; k < UP_CH * UP_CW; k += 256) {
        const int qx = qxa + k % UP_CW, qy = qya + k / UP_CW;
        float4 r = make_float4(0.f, 0.f, 0.f, __int_as_float(0x7fc00000));
        float g = 0.f;
        if (qx >= 0 && qx < L.w && qy >= 0 && qy < L.h) {
            r = plane_record(low[qy * L.w + qx], pixel_ray(L, qx, qy));
            g = (float)P.gray[L.goff + (size_t)qy * L.w + qx] * (1.0f / 255.0f);
        }
        rec[k] = r;
        lgray[k] = g;
    }
    __syncthreads();
    const int x = px0 + (threadIdx.x & 63), y = py0 + (threadIdx.x >> 6);
    if (x >= it.w || y >= it.h) return;
    const f3 rp = pixel_ray(R, x, y);
    const float ip = (float)P.gray[R.goff + (size_t)y * R.w + x] * (1.0f / 255.0f);
    const float px = ((float)x - 0.5f) * 0.5f, py = ((float)y - 0.5f) * 0.5f;
    const int qx0 = (int)roundf(px), qy0 = (int)roundf(py);
    float sw = 0.f, sd = 0.f, snx = 0.f, sny = 0.f, snz = 0.f;
    float nd = 1e30f, ndep = 0.f;
    f3 nn = mk3(0.f, 0.f, 0.f);
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
        const int qy = qy0 + dy;
        const int row = (qy - qya) * UP_CW - qxa;
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const int qx = qx0 + dx;
            const float4 r = rec[row + qx];
            const f3 nq = xyz(r);
            const float den = dot3(nq, rp);
            const float d = r.w / den;
            if (!(den < -1e-6f) || !(d > 0.f) || !isfinite(d)) continue;
            const float ds = ((float)qx - px) * ((float)qx - px) + ((float)qy - py) * ((float)qy - py);
            if (ds < nd) {
                nd = ds;
                ndep = d;
                nn = nq;
            }
            const float iq = lgray[row + qx];
            const float wgt = c3d_expf(-ds * P.up_inv2s - (ip - iq) * (ip - iq) * P.up_inv2c);
            sw += wgt;
            sd += wgt * d;
            snx += wgt * nq.x;
            sny += wgt * nq.y;
            snz += wgt * nq.z;
        }
    }
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    const bool weighted = sw > 1e-12f;
    if (weighted || nd < 1e29f) {
        const float d = weighted ? sd / sw : ndep;
        f3 nv = weighted ? mk3(snx, sny, snz) : nn;
        if (dot3(nv, nv) > 1e-24f) {
            nv = nrm3(nv);
            if (dot3(nv, rp) >= 0.f) nv = mk3(-nv.x, -nv.y, -nv.z);
            o = mk4(nv, d);
        }
    }
    P.plane[it.off + y * it.w + x] = o;
}

// ================================================================ 5 x 5 median plane filter
// Output rule of pm_median (patchmatch.cu:803-849, upsample.rs:177-260): a pixel with a positive
// depth takes, among the neighbours of its 5 x 5 block (clipped to the image) whose plane transfers
// to a positive finite depth on its ray, the lower median of those depths, ties ordered by pixel
// index; its output is that neighbour's normal with the transferred depth. Other pixels keep their
// plane.
//
// The 64 x 4 tile and its 2-pixel halo (68 x 8 records) are staged in LDS. Each pixel then holds
// its 25 candidate depths in registers (a rejected one as +inf) and picks the median by rank:
// candidate k has rank #{j > k : d_j < d_k} + #{j < k : d_j <= d_k} (slot order is pixel-index
// order, so this is the index tie-break), and the candidate of rank (count - 1) / 2 is the median.
// No data-dependent branches or register indexing: every lane does the same 600 comparisons.
constexpr int MED_CW = PIX_TW + 4;
constexpr int MED_CH = PIX_TH + 4;

__global__ __launch_bounds__(256) void k_median(KParams P) {
    __shared__ float4 rec[MED_CH * MED_CW];
    BatchItem it;
    int px0, py0;
    if (!find_tile(P, xcd_tile(blockIdx.x, P.ntiles), PIX_TW, PIX_TH, it, px0, py0)) return;  // uniform per workgroup
    const LevelDesc R = P.lv[it.view * P.nlevels + P.level];
    const float4* in = P.prior + it.off;
    const int w = it.w, h = it.h;
    const int xa = px0 - 2, ya = py0 - 2;
    for (int k = threadIdx.x; k < MED_CH * MED_CW; k += 256) {
        const int qx = xa + k % MED_CW, qy = ya + k / MED_CW;
        float4 r = make_float4(0.f, 0.f, 0.f, __int_as_float(0x7fc00000));
        if (qx >= 0 && qx < w && qy >= 0 && qy < h) r = plane_record(in[qy * w + qx], pixel_ray(R, qx, qy));
        rec[k] = r;
    }
    __syncthreads();
    const int x = px0 + (threadIdx.x & 63), y = py0 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const float4 self = in[y * w + x];
    float4 out = self;
    if (self.w > 0.f) {
        const f3 rp = pixel_ray(R, x, y);
        const int base = (y - ya - 2) * MED_CW + (x - xa - 2);
        float d[25];
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < 25; ++k) {
            const float4 r = rec[base + (k / 5) * MED_CW + k % 5];
            const float den = dot3(xyz(r), rp);
            const float t = r.w / den;
            const bool ok = den < -1e-6f && t > 0.f && isfinite(t);
            d[k] = ok ? t : __int_as_float(0x7f800000);
            cnt += ok ? 1 : 0;
        }
        if (cnt > 0) {
            const int m = (cnt - 1) / 2;
            int pick = 0;
#pragma unroll
            for (int k = 0; k < 25; ++k) {
                int rank = 0;
#pragma unroll
                for (int j = 0; j < 25; ++j) {
                    if (j < k) rank += d[j] <= d[k] ? 1 : 0;
                    if (j > k) rank += d[j] < d[k] ? 1 : 0;
                }
                pick = (rank == m && d[k] < __int_as_float(0x7f800000)) ? k : pick;
            }
            float dm = d[0];
#pragma unroll
            for (int k = 1; k < 25; ++k) dm = pick == k ? d[k] : dm;
            const float4 r = rec[base + (pick / 5) * MED_CW + pick % 5];
            out = make_float4(r.x, r.y, r.z, dm);
        }
    }
    P.plane[it.off + y * w + x] = out;
}

// ================================================================ small device-side helpers

// Quad tables of one view-level: entry (x0, y0), x0 in [-2, w], y0 in [-2, h].
__global__ void k_quad_build(const uint8_t* g, int w, int h, uint32_t* q) {
    int i = blockIdx.x * 256 + threadIdx.x;
    int qp = w + 3;
    if (i >= qp * (h + 3)) return;
    int y0 = i / qp - 2, x0 = i % qp - 2;
    auto at = [&](int x, int y) -> uint32_t {
        return (x >= 0 && y >= 0 && x < w && y < h) ? (uint32_t)g[(size_t)y * w + x] : 0u;
    };
    q[i] = at(x0, y0) | (at(x0 + 1, y0) << 8) | (at(x0, y0 + 1) << 16) | (at(x0 + 1, y0 + 1) << 24);
}

// depth channel of the planes -> snapshot buffer (same pixel layout)
__global__ void k_extract_depth(const float4* plane, float* out, long long n) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = plane[i].w;
}

// read-out threshold: depth/normal kept where depth > 0 and count >= need (densify.rs:446-454)
__global__ __launch_bounds__(256) void k_threshold(KParams P, const uint8_t* need_per_item, float4* out, long long npix) {
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= npix) return;
    BatchItem it;
    int item, pix;
    if (!locate_pixel(P, g, it, item, pix)) return;
    const long long p = it.off + pix;
    float4 h = P.plane[p];
    if (!(h.w > 0.f && P.count[p] >= need_per_item[item])) h = make_float4(0.f, 0.f, 0.f, 0.f);
    out[p] = h;
}

// detail restorer merge (densify.rs:396-409): the photometric plane replaces the upsampled one where
// the upsampled depth is not positive or c_init - c_photo > xi; the merged state carries cost 2 (the
// upsampled ViewState is ViewState::new, kernel.rs:154-159).
__global__ void k_restore(float4* plane, const float4* up, const float* c_init, float* cost, float xi, long long n) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 u = up[i];
    if (!(u.w <= 0.f || c_init[i] - cost[i] > xi)) plane[i] = u;
    cost[i] = 2.0f;
}

// one step of the gray pyramid (build_pyramid and blur3, crates/dense/src/upsample.rs:45-99): the
// 3-tap Gaussian (edge clamp, rows first, then columns) at the four pixels of the 2x2 cell, their
// mean, rounded half away from zero and clamped to 8 bits. Every sum runs in the reference's order.
__global__ void k_pyramid_step(const uint8_t* src, int w, int h, uint8_t* dst, float g0, float g1, float g2) {
    const int nw = w / 2, nh = h / 2;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nw * nh) return;
    const int x = i % nw, y = i / nw;
    const float g[3] = {g0, g1, g2};
    float b[4];
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const int X = 2 * x + dx, Y = 2 * y + dy;
            const int xa = X - 1 < 0 ? 0 : X - 1, xb = X + 1 > w - 1 ? w - 1 : X + 1;
            float a = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                int yy = Y + k - 1;
                yy = yy < 0 ? 0 : (yy > h - 1 ? h - 1 : yy);
                const uint8_t* r = src + (size_t)yy * w;
                float t = 0.0f;
                t += g0 * (float)r[xa];
                t += g1 * (float)r[X];
                t += g2 * (float)r[xb];
                a += g[k] * t;
            }
            b[dy * 2 + dx] = a;
        }
    const float v = 0.25f * (b[0] + b[1] + b[2] + b[3]);
    const float r = roundf(v);
    dst[i] = (uint8_t)(r < 0.f ? 0.f : (r > 255.f ? 255.f : r));
}

// ================================================================ launchers

bool half_shape_ok(int rows, int tpp) { return ((rows == 1 || rows == 2) && tpp == 1) || (rows == 4 && (tpp == 1 || tpp == 2)); }

int half_tiles(const BatchItem* items, int n, int rows, int tpp) {
    const int span = HALF_SPAN / tpp;
    int t = 0;
    for (int i = 0; i < n; ++i) t += ((items[i].w + span - 1) / span) * ((items[i].h + rows - 1) / rows);
    return t;
}
int pix_tiles(const BatchItem* items, int n) {
    int t = 0;
    for (int i = 0; i < n; ++i) t += ((items[i].w + PIX_TW - 1) / PIX_TW) * ((items[i].h + PIX_TH - 1) / PIX_TH);
    return t;
}

// dynamic LDS bytes of the any-radius (WR = 0) reference tiles
static size_t half_tile_bytes(int wr, int rows, int tpp) { return (size_t)(HALF_SPAN / tpp + 2 * wr) * (rows + 2 * wr) * sizeof(float); }
static size_t pix_tile_bytes(int wr) { return (size_t)(PIX_TW + 2 * wr) * (PIX_TH + 2 * wr) * sizeof(float); }

template <int WR, int WSTEP, int ROWS, int TPP>
static void launch_half_rows(const KParams& P, size_t lds, hipStream_t s) {
    if (P.geometric) hipLaunchKernelGGL((k_half<WR, WSTEP, true, ROWS, TPP>), dim3(P.ntiles), dim3(ROWS * 64), lds, s, P);
    else hipLaunchKernelGGL((k_half<WR, WSTEP, false, ROWS, TPP>), dim3(P.ntiles), dim3(ROWS * 64), lds, s, P);
}

// P.ntiles must count the tiles of the shape (half_tiles(..., rows, tpp)); the compiled shapes are
// (rows, lanes per pixel) = (1, 1), (2, 1), (4, 1) and (4, 2)
template <int WR, int WSTEP>
static hipError_t launch_half(const KParams& P0, int rows, int tpp, hipStream_t s) {
    if (!half_shape_ok(rows, tpp)) return hipErrorInvalidValue;
    KParams P = P0;
    P.wv_stride = rows * 64;
    size_t lds = (size_t)P.wv_stride * (P.max_nsrc > 0 ? P.max_nsrc : 1) * sizeof(float);
    if (WR == 0) lds += half_tile_bytes(P.wr, rows, tpp);
    if (rows == 1) launch_half_rows<WR, WSTEP, 1, 1>(P, lds, s);
    else if (rows == 2) launch_half_rows<WR, WSTEP, 2, 1>(P, lds, s);
    else if (tpp == 1) launch_half_rows<WR, WSTEP, 4, 1>(P, lds, s);
    else launch_half_rows<WR, WSTEP, 4, 2>(P, lds, s);
    return hipGetLastError();
}
template <int WR, int WSTEP>
static hipError_t launch_eval(const KParams& P, int mode, hipStream_t s) {
    hipLaunchKernelGGL((k_eval<WR, WSTEP, false>), dim3(P.ntiles), dim3(256), WR == 0 ? pix_tile_bytes(P.wr) : 0, s, P, mode);
    return hipGetLastError();
}
// P.ntiles must count 128 x 4 spans (half_tiles(..., 4, 1)); P.color selects the pixels
template <int WR, int WSTEP>
static hipError_t launch_eval_color(const KParams& P, hipStream_t s) {
    hipLaunchKernelGGL((k_eval<WR, WSTEP, true>), dim3(P.ntiles), dim3(256), WR == 0 ? half_tile_bytes(P.wr, PIX_TH, 1) : 0, s, P, 0);
    return hipGetLastError();
}
template <int WR, int WSTEP>
static hipError_t launch_filter(const KParams& P, hipStream_t s) {
    hipLaunchKernelGGL((k_filter<WR, WSTEP>), dim3(P.ntiles), dim3(256), WR == 0 ? pix_tile_bytes(P.wr) : 0, s, P);
    return hipGetLastError();
}
template <int WR, int WSTEP>
static hipError_t launch_pair_costs(const KParams& P, int view, int slot, int n, const int* xy, const float4* planes, float* out,
                                    hipStream_t s) {
    hipLaunchKernelGGL((k_pair_costs<WR, WSTEP>), dim3((n + 255) / 256), dim3(256), 0, s, P, view, slot, n, xy, planes, out);
    return hipGetLastError();
}
template <int WR, int WSTEP>
static const void* half_symbol() {
    return reinterpret_cast<const void*>(&k_half<WR, WSTEP, false, 4, 1>);
}

template <int WR, int WSTEP>
static constexpr KernelSet make_set() {
    return KernelSet{WR, WSTEP, &launch_half<WR, WSTEP>, &launch_eval<WR, WSTEP>, &launch_eval_color<WR, WSTEP>,
                     &launch_filter<WR, WSTEP>, &launch_pair_costs<WR, WSTEP>, &half_symbol<WR, WSTEP>};
}

// Window variants compiled into the library (radius, step): the fast profile (5, 2) and the
// variants of the reference's window report (patchmatch_gpu.rs:164) whose window holds at most
// 49 samples, with the bilateral weights in registers; every other radius up to 20 (the quality
// profile's (5, 1) among them) runs the any-radius kernels (wr = 0 below) of its step.
static const KernelSet kSets[] = {make_set<5, 2>(), make_set<4, 2>(), make_set<3, 2>(), make_set<2, 2>(), make_set<3, 1>(),
                                  make_set<2, 1>(), make_set<1, 1>()};
static const KernelSet kAny[] = {make_set<0, 1>(), make_set<0, 2>()};

const KernelSet* kernel_set(int wr, int wstep, bool any_radius) {
    if (wr < 1 || wr > MAX_WR || wstep < 1 || wstep > 2) return nullptr;
    if (!any_radius)
        for (const KernelSet& k : kSets)
            if (k.wr == wr && k.wstep == wstep) return &k;
    return &kAny[wstep - 1];
}

// The per-sample reciprocal of the window loops, applied to n given values (c3d_pm_sample_reciprocal).
__global__ __launch_bounds__(256) void k_sample_reciprocal(const float* z, float* r, long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) r[i] = rcp_sample(z[i]);
}

hipError_t launch_sample_reciprocal(const float* z, float* r, long long n, hipStream_t s) {
    hipLaunchKernelGGL(k_sample_reciprocal, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, z, r, n);
    return hipGetLastError();
}

// a depth array and a 3-float normal array merged into planes (nx, ny, nz, d) (state uploads)
__global__ __launch_bounds__(256) void k_merge_planes(const float* depth, const float* normal, float4* plane, long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) plane[i] = make_float4(normal[3 * i], normal[3 * i + 1], normal[3 * i + 2], depth[i]);
}

hipError_t launch_merge_planes(const float* depth, const float* normal, float4* plane, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_merge_planes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, depth, normal, plane, n);
    return hipGetLastError();
}

// planes (nx, ny, nz, d) split into a depth array and a 3-float normal array (result downloads)
__global__ __launch_bounds__(256) void k_split_planes(const float4* plane, float* depth, float* normal, long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const float4 q = plane[i];
        depth[i] = q.w;
        normal[3 * i] = q.x;
        normal[3 * i + 1] = q.y;
        normal[3 * i + 2] = q.z;
    }
}

hipError_t launch_split_planes(const float4* plane, float* depth, float* normal, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_split_planes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, plane, depth, normal, n);
    return hipGetLastError();
}

// npix: the pixels of the batch (the end of its last item)
hipError_t launch_eval_init(const KParams& P, long long npix, hipStream_t s) {
    if (npix <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_eval_init, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, P, npix);
    return hipGetLastError();
}

hipError_t launch_upsample(const KParams& P, hipStream_t s) {
    hipLaunchKernelGGL(k_upsample, dim3(P.ntiles), dim3(256), 0, s, P);
    return hipGetLastError();
}
hipError_t launch_median(const KParams& P, hipStream_t s) {
    hipLaunchKernelGGL(k_median, dim3(P.ntiles), dim3(256), 0, s, P);
    return hipGetLastError();
}
hipError_t launch_quad_build(const uint8_t* g, int w, int h, uint32_t* q, hipStream_t s) {
    int n = (w + 3) * (h + 3);
    hipLaunchKernelGGL(k_quad_build, dim3((n + 255) / 256), dim3(256), 0, s, g, w, h, q);
    return hipGetLastError();
}
hipError_t launch_extract_depth(const float4* plane, float* out, long long n, hipStream_t s) {
    hipLaunchKernelGGL(k_extract_depth, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, plane, out, n);
    return hipGetLastError();
}
hipError_t launch_restore(float4* plane, const float4* up, const float* c_init, float* cost, float xi, long long n,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_restore, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, plane, up, c_init, cost, xi, n);
    return hipGetLastError();
}
hipError_t launch_pyramid_step(const uint8_t* src, int w, int h, uint8_t* dst, const float g[3], hipStream_t s) {
    const int n = (w / 2) * (h / 2);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pyramid_step, dim3((n + 255) / 256), dim3(256), 0, s, src, w, h, dst, g[0], g[1], g[2]);
    return hipGetLastError();
}
// Tile offsets of the batch items for tiles of tw x th pixels: pre[i] = tiles of the items before
// i, pre[n] = all tiles. One lane sums the counts in item order (a few hundred items at most).
__global__ void k_tile_prefix(const BatchItem* batch, int n, int tw, int th, int* pre) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int acc = 0;
    for (int i = 0; i < n; ++i) {
        const BatchItem b = batch[i];
        pre[i] = acc;
        acc += ((b.w + tw - 1) / tw) * ((b.h + th - 1) / th);
    }
    pre[n] = acc;
}

// pre: n + 1 ints; the half step's tiles for (rows, lanes per pixel) are HALF_SPAN / tpp x rows
hipError_t launch_half_tile_prefix(const BatchItem* batch, int n, int rows, int tpp, int* pre, hipStream_t s) {
    hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(64), 0, s, batch, n, HALF_SPAN / tpp, rows, pre);
    return hipGetLastError();
}
// the 64 x 4 tiles of k_eval (all pixels) and k_filter (pix_tiles)
hipError_t launch_pix_tile_prefix(const BatchItem* batch, int n, int* pre, hipStream_t s) {
    hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(64), 0, s, batch, n, PIX_TW, PIX_TH, pre);
    return hipGetLastError();
}
// the 128 x 4 spans of k_eval over the pixels of one colour (half_tiles(..., 4, 1))
hipError_t launch_span_tile_prefix(const BatchItem* batch, int n, int* pre, hipStream_t s) {
    hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(64), 0, s, batch, n, HALF_SPAN, PIX_TH, pre);
    return hipGetLastError();
}

// npix: the pixels of the batch (the end of its last item)
hipError_t launch_threshold(const KParams& P, const uint8_t* need, float4* out, long long npix, hipStream_t s) {
    if (npix <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_threshold, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, P, need, out, npix);
    return hipGetLastError();
}

}  // namespace c3d
