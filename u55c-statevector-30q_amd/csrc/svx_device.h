// Device helpers shared by the kernel translation units of libsvx (gfx950): the complex pair type,
// bit deposits, the padded LDS tile layout, 16-byte streaming HBM access, the HBM <-> LDS tile
// moves, diagonal tables on the LDS tile and the workgroup-to-tile order.
#pragma once
#include "svx_internal.h"

namespace svx {

template <typename T>
struct C2 {
    T x, y;
};

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

__device__ __forceinline__ uint32_t lpad(uint32_t e) { return e + (e >> 5); }

__device__ __forceinline__ uint64_t ins0_64(uint64_t x, int p) {
    const uint64_t lo = x & ((uint64_t(1) << p) - 1);
    return ((x >> p) << (p + 1)) | lo;
}
__device__ __forceinline__ uint32_t ins0_32(uint32_t x, int p) {
    const uint32_t lo = x & ((1u << p) - 1u);
    return ((x >> p) << (p + 1)) | lo;
}

// ------------------------------------------------------------------------------------------ pass
// HBM <-> LDS for one workgroup tile.  Element e of the workgroup tile: QSV2 tile u = e >> k,
// tile index j = e & (2^k - 1); its state index is tb[u] | sbo[j >> (k-m)] | (j & low mask).
#ifndef SVX_TILE_NT
#define SVX_TILE_NT 1
#endif
typedef float f4v __attribute__((ext_vector_type(4)));
// 16-byte state access that streams past the caches (the state is touched once per sweep)
__device__ __forceinline__ float4 ld_stream(const float *p) {
#if SVX_TILE_NT
    const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(p));
    return make_float4(v[0], v[1], v[2], v[3]);
#else
    return *reinterpret_cast<const float4 *>(p);
#endif
}
__device__ __forceinline__ void st_stream(float *p, float4 v) {
#if SVX_TILE_NT
    __builtin_nontemporal_store(f4v{v.x, v.y, v.z, v.w}, reinterpret_cast<f4v *>(p));
#else
    *reinterpret_cast<float4 *>(p) = v;
#endif
}

template <typename T, int K, bool LOAD, int NT = kNT>
__device__ __forceinline__ void move_tile(C2<T> *__restrict__ st, C2<T> *tile, const uint64_t *sbo,
                                          const uint64_t *tb, int k, int lowb, uint32_t tid, int unscale = 0) {
    constexpr uint32_t NE = 1u << K;
    const uint32_t lowmask = (1u << lowb) - 1u, tmask = (1u << k) - 1u;
    if (sizeof(T) == 4 && lowb >= 1) {
        // two adjacent amplitudes (16 bytes) per lane and access
        constexpr int IT = NE / (2 * NT);
        constexpr int BATCH = IT < 8 ? IT : 8;
#pragma unroll 1
        for (int b0 = 0; b0 < IT; b0 += BATCH) {
            float4 v[BATCH];
            uint64_t g[BATCH];
#pragma unroll
            for (int i = 0; i < BATCH; ++i) {
                const uint32_t e = 2u * ((b0 + i) * NT + tid);
                const uint32_t j = e & tmask, u = e >> k;
                g[i] = tb[u] | sbo[j >> lowb] | (j & lowmask);
            }
            if (LOAD) {
#pragma unroll
                for (int i = 0; i < BATCH; ++i)
                    v[i] = ld_stream(reinterpret_cast<const float *>(st) + 2 * g[i]);
#pragma unroll
                for (int i = 0; i < BATCH; ++i) {
                    const uint32_t e = 2u * ((b0 + i) * NT + tid), pe = lpad(e);
                    tile[pe] = C2<T>{T(v[i].x), T(v[i].y)};
                    tile[pe + 1] = C2<T>{T(v[i].z), T(v[i].w)};
                }
            } else {
#pragma unroll
                for (int i = 0; i < BATCH; ++i) {
                    const uint32_t e = 2u * ((b0 + i) * NT + tid), pe = lpad(e);
                    const C2<T> p0 = tile[pe], p1 = tile[pe + 1];
                    v[i] = make_float4(float(p0.x), float(p0.y), float(p1.x), float(p1.y));
                    if (unscale) {
                        v[i].x = ldexpf(v[i].x, unscale);
                        v[i].y = ldexpf(v[i].y, unscale);
                        v[i].z = ldexpf(v[i].z, unscale);
                        v[i].w = ldexpf(v[i].w, unscale);
                    }
                }
#pragma unroll
                for (int i = 0; i < BATCH; ++i)
                    st_stream(reinterpret_cast<float *>(st) + 2 * g[i], v[i]);
            }
        }
    } else {
        constexpr int IT = NE / NT;
        constexpr int BATCH = IT < 8 ? IT : 8;
#pragma unroll 1
        for (int b0 = 0; b0 < IT; b0 += BATCH) {
            C2<T> v[BATCH];
            uint64_t g[BATCH];
#pragma unroll
            for (int i = 0; i < BATCH; ++i) {
                const uint32_t e = (b0 + i) * NT + tid;
                const uint32_t j = e & tmask, u = e >> k;
                g[i] = tb[u] | sbo[j >> lowb] | (j & lowmask);
            }
            if (LOAD) {
#pragma unroll
                for (int i = 0; i < BATCH; ++i) v[i] = st[g[i]];
#pragma unroll
                for (int i = 0; i < BATCH; ++i) tile[lpad((b0 + i) * NT + tid)] = v[i];
            } else {
#pragma unroll
                for (int i = 0; i < BATCH; ++i) {
                    v[i] = tile[lpad((b0 + i) * NT + tid)];
                    if (unscale) {
                        v[i].x = T(ldexpf(float(v[i].x), unscale));
                        v[i].y = T(ldexpf(float(v[i].y), unscale));
                    }
                }
#pragma unroll
                for (int i = 0; i < BATCH; ++i) st[g[i]] = v[i];
            }
        }
    }
}

// |x| as an ordered integer (non-negative floats order like their bit patterns)
__device__ __forceinline__ uint32_t absbits(float x) { return __float_as_uint(x) & 0x7fffffffu; }

// Diagonal table on the padded LDS tile: element e times entry e mod 2^k (its QSV2 tile index) of
// the 2^k complex64 phases at tab; consecutive threads take consecutive elements, so both the LDS
// accesses and the table reads (cached: every workgroup reads the same table) are contiguous.
template <int K, int NT>
__device__ __forceinline__ void diag_table(C2<float> *tile, const float *__restrict__ tab, int k, uint32_t tid) {
    constexpr uint32_t NE = 1u << K;
    const uint32_t km = (1u << k) - 1u;
    const float2 *d = reinterpret_cast<const float2 *>(tab);
#pragma unroll 4
    for (uint32_t e = tid; e < NE; e += NT) {
        const float2 w = d[e & km];
        const C2<float> a = tile[lpad(e)];
        tile[lpad(e)] = C2<float>{fmaf(a.x, w.x, -a.y * w.y), fmaf(a.x, w.y, a.y * w.x)};
    }
}

// Order in which the workgroups of a pass take the 2^tbits tiles (tbits = nbits - K).  Workgroups are
// dispatched round-robin over the 8 XCDs; the default (-2) gives each XCD one contiguous eighth of the
// tiles, which measured 3-18% faster than the identity order on light 30-qubit passes.
__device__ __forceinline__ uint64_t tile_order(uint64_t b, const PassDev *pd) {
    const int tbits = pd->nbits - pd->K, r = pd->torder;
    if (r == 0 || tbits <= 1) return b;
    if (r == -2) {   // XCD-contiguous: workgroups go round-robin over 8 XCDs; XCD x takes tiles x T/8 ..
        if (tbits < 3) return b;
        return ((b & 7u) << (tbits - 3)) | (b >> 3);
    }
    if (r < 0) return __builtin_bitreverse64(b) >> (64 - tbits);
    const int rr = r % tbits;
    const uint64_t mask = (uint64_t(1) << tbits) - 1;
    return ((b << rr) | (b >> (tbits - rr))) & mask;
}

// The LDS tile layout is a template choice: padded (byte address 8 lpad(e), additive over disjoint
// bits: the single-buffer kernel) or XOR-swizzled (swz(e), XOR-linear: the double-buffered kernel).
typedef float f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint32_t swz(uint32_t e) {
    return (((e >> 1) ^ ((e >> 5) & 15u)) << 4) | ((e & 1u) << 3);
}

template <bool SWZ>
struct TileAddr {
    __device__ __forceinline__ static uint32_t of(uint32_t e) { return SWZ ? swz(e) : 8u * lpad(e); }
    __device__ __forceinline__ static uint32_t comb(uint32_t a, uint32_t b) { return SWZ ? (a ^ b) : (a + b); }
};

// byte offsets of the 2^F components of one column, folded from the per-bit offsets
template <int F, bool SWZ>
struct ColOff {
    uint32_t lq[F];
    __device__ __forceinline__ void init(const GroupDev &G) {
#pragma unroll
        for (int i = 0; i < F; ++i) lq[i] = TileAddr<SWZ>::of(1u << G.dq[i]);
    }
    __device__ __forceinline__ uint32_t operator()(uint32_t c) const {
        uint32_t o = 0;
#pragma unroll
        for (int i = 0; i < F; ++i)
            if ((c >> i) & 1u) o = TileAddr<SWZ>::comb(o, lq[i]);
        return o;
    }
};

__device__ __forceinline__ C2<float> ldc(const unsigned char *buf, uint32_t a) {
    return *reinterpret_cast<const C2<float> *>(buf + a);
}
__device__ __forceinline__ void stc(unsigned char *buf, uint32_t a, C2<float> v) {
    *reinterpret_cast<C2<float> *>(buf + a) = v;
}
// re and im of one complex64 tile element as two 32-bit LDS stores (paired into one ds_write2 with
// no register moves, where a 64-bit store would need the two values in adjacent registers)
__device__ __forceinline__ void stc2(unsigned char *buf, uint32_t a, float re, float im) {
    float *p = reinterpret_cast<float *>(buf + a);
    p[0] = re;
    p[1] = im;
}


// FP32 matrix cores (exact for 0 / +-1 matrices): v_mfma_f32_32x32x2f32, half h = l / 32 feeds K-rows
// 2s + h; W arrives in A-operand order wd[(rt 2^F + s) 64 + l] = W[32 rt + l % 32][2 s + l / 32].
// NWAVE wavefronts (wt = 0 .. 64 NWAVE - 1) share the columns.
template <int K, int F, int NWAVE, bool SWZ>
__device__ __forceinline__ void dense_apply(unsigned char *buf, const GroupDev &G, const float *__restrict__ wd,
                                            uint32_t wt) {
    constexpr int NRT = F == 5 ? 2 : 1;            // 32-row output tiles
    constexpr int NS = 1 << F;                     // K-steps of 2 real rows
    constexpr int NX = NS / 2;                     // complex inputs per lane
    constexpr uint32_t NB = (1u << (K - F)) / 32;  // 32-column batches per tile
    const uint32_t lane = wt & 63u, h = lane >> 5, j = lane & 31u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(wt >> 6);
    const float *w = wd + G.w_off;
    float wa[NRT][NS];
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int s = 0; s < NS; ++s) wa[rt][s] = w[(rt * NS + s) * 64 + lane];
    ColOff<F, SWZ> coff;
    coff.init(G);
    const uint32_t hoff = h ? coff.lq[0] : 0u;
    for (uint32_t b = wave; b < NB; b += NWAVE) {
        uint32_t base = b * 32u + j;
#pragma unroll
        for (int i = 0; i < F; ++i) base = ins0_32(base, G.dqs[i]);
        const uint32_t pb = TileAddr<SWZ>::of(base);
        C2<float> x[NX];
#pragma unroll
        for (int s = 0; s < NX; ++s) x[s] = ldc(buf, TileAddr<SWZ>::comb(pb, TileAddr<SWZ>::comb(coff(2u * s), hoff)));
        f16v acc[NRT];
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt) {
            acc[rt] = f16v{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const float bv = s < NX ? x[s].x : x[s - NX].y;
                acc[rt] = __builtin_amdgcn_mfma_f32_32x32x2f32(wa[rt][s], bv, acc[rt], 0, 0, 0);
            }
        }
        if (F == 4) {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const uint32_t c = 8u * (v / 4) + 4u * h + (v % 4);
                stc2(buf, TileAddr<SWZ>::comb(pb, coff(c)), acc[0][v], acc[0][v + 8]);
            }
        } else {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const uint32_t c = 8u * (v / 4) + 4u * h + (v % 4);
                stc2(buf, TileAddr<SWZ>::comb(pb, coff(c)), acc[0][v], acc[NRT - 1][v]);
            }
        }
    }
}

typedef float f4acc __attribute__((ext_vector_type(4)));

// FP32 matrix cores, 4-qubit units as three real 16 x 16 products (Gauss's complex product): with
// W = Wr + i Wi and a column x = p + i q, k1 = (Wr + Wi) p, Re y = k1 - Wi (p + q),
// Im y = k1 + Wr (q - p) -- 48 multiply-adds per amplitude instead of 64.  v_mfma_f32_16x16x4f32:
// lane l holds A[l % 16][l / 16] and B[l / 16][l % 16], and D[4 (l / 16) + v][l % 16] in register v.
// K step t, lane quarter g = l / 16 carries component 4 g + t, so a lane reads components
// 4 g .. 4 g + 3 of column l % 16 and writes the same four outputs (no other lane touches them).
// Operands: wd[(s 4 + t) 64 + l] = W_s[l % 16][4 (l / 16) + t] with W_0 = Wr + Wi, W_1 = -Wi, W_2 = Wr.
// Each wavefront step carries up to two 16-column blocks, independent MFMA chains.
// Addresses: column j of block blk sits on the tile element with j's four bits on the unit's lane
// bits (GroupDev::jmask, chosen by the host so that 16 lanes hit distinct LDS banks) and blk's bits
// on the remaining free bits (ascending).  The deposit of disjoint parts is the sum (padded layout)
// or XOR (swizzled) of their deposits, so a lane's part (j, and its components' offsets) is formed
// once per unit and the block part (wavefront-uniform) in scalar registers per block.
// the lane's twelve A-operand values of a 4-qubit three-product unit (loaded ahead of the unit by
// callers that can, so the loads overlap a barrier or the tile load)
struct Gauss4W {
    float w[12];
};
__device__ __forceinline__ void gauss4_weights(const float *__restrict__ w, uint32_t lane, Gauss4W &out) {
#pragma unroll
    for (int i = 0; i < 12; ++i) out.w[i] = w[i * 64 + (lane & 63u)];
}
template <int K, int NWAVE, bool SWZ>
__device__ __forceinline__ void dense_gauss4(unsigned char *buf, const GroupDev &G, const Gauss4W &gw, uint32_t wt) {
    constexpr uint32_t NB = (1u << (K - 4)) / 16;   // 16-column blocks per tile
    const uint32_t lane = wt & 63u, g = lane >> 4, j = lane & 15u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(wt >> 6);
    float w0[4], w1[4], w2[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        w0[t] = gw.w[t];
        w1[t] = gw.w[4 + t];
        w2[t] = gw.w[8 + t];
    }
    // lane bits and the block bits' positions (wavefront-uniform)
    const uint32_t jm = __builtin_amdgcn_readfirstlane((uint32_t)G.jmask);
    uint32_t rest = ((1u << K) - 1u) & ~jm;
#pragma unroll
    for (int i = 0; i < 4; ++i) rest &= ~(1u << G.dqs[i]);
    constexpr int NBB = K - 8;   // block-index bits
    uint32_t bp[NBB > 0 ? NBB : 1];
#pragma unroll
    for (int i = 0; i < NBB; ++i) {
        bp[i] = __builtin_ctz(rest);
        rest &= rest - 1u;
    }
    auto dep = [&](uint32_t blk) {
        uint32_t e = 0;
#pragma unroll
        for (int i = 0; i < NBB; ++i) e |= ((blk >> i) & 1u) << bp[i];
        return e;
    };
    uint32_t ej = 0;
    {
        uint32_t m = jm;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ej |= ((j >> i) & 1u) << __builtin_ctz(m);
            m &= m - 1u;
        }
    }
    ColOff<4, SWZ> coff;
    coff.init(G);
    const uint32_t lane_off = TileAddr<SWZ>::of(ej);
    uint32_t co[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) co[t] = TileAddr<SWZ>::comb(lane_off, coff(4u * g + t));
    // BPS blocks per wavefront step (independent MFMA chains): the tile's blocks split evenly over
    // the wavefronts, at most two per step (four measured slower)
    constexpr uint32_t PER = NB / NWAVE > 0 ? NB / NWAVE : 1;
    constexpr uint32_t BPS = PER < 2 ? PER : 2;
    for (uint32_t b = wave; b < NB; b += BPS * NWAVE) {
        uint32_t ub[BPS];
        bool ok[BPS];
        C2<float> x[BPS][4];
#pragma unroll
        for (uint32_t h = 0; h < BPS; ++h) {
            const uint32_t blk = b + h * NWAVE;
            ok[h] = blk < NB;   // uniform
            ub[h] = TileAddr<SWZ>::of(dep(ok[h] ? blk : b));
#pragma unroll
            for (int t = 0; t < 4; ++t) x[h][t] = ldc(buf, TileAddr<SWZ>::comb(co[t], ub[h]));
        }
        f4acc k1[BPS], re[BPS], im[BPS];
#pragma unroll
        for (uint32_t h = 0; h < BPS; ++h) k1[h] = f4acc{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (uint32_t h = 0; h < BPS; ++h) k1[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(w0[t], x[h][t].x, k1[h], 0, 0, 0);
#pragma unroll
        for (uint32_t h = 0; h < BPS; ++h) {
            re[h] = k1[h];
            im[h] = k1[h];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (uint32_t h = 0; h < BPS; ++h) {
                re[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(w1[t], x[h][t].x + x[h][t].y, re[h], 0, 0, 0);
                im[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(w2[t], x[h][t].y - x[h][t].x, im[h], 0, 0, 0);
            }
#pragma unroll
        for (uint32_t h = 0; h < BPS; ++h)
            if (ok[h])
#pragma unroll
                for (int v = 0; v < 4; ++v) stc2(buf, TileAddr<SWZ>::comb(co[v], ub[h]), re[h][v], im[h][v]);
    }
}


// FP32 dense groups (type 1: 32x32x2 real form; type 5: 4-qubit three-product form) by NWAVE
// wavefronts, for kernels without the split-FP16 paths
template <int K, int NWAVE, bool SWZ>
__device__ __forceinline__ void dense_group_f32(unsigned char *buf, const GroupDev &G, const float *__restrict__ wd,
                                                uint32_t wt) {
    if (G.type == 5) {
        Gauss4W gw;
        gauss4_weights(wd + G.w_off, wt, gw);
        dense_gauss4<K, NWAVE, SWZ>(buf, G, gw, wt);
    }
    else if (G.f == 5) dense_apply<K, 5, NWAVE, SWZ>(buf, G, wd, wt);
    else dense_apply<K, 4, NWAVE, SWZ>(buf, G, wd, wt);
}

}  // namespace svx
