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

}  // namespace svx
