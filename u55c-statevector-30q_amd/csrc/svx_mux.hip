// The FP32 MUX2 pass kernel of libsvx (gfx950): complex64 passes of MUX2 ops and diagonal tables.
#include "svx_device.h"

namespace svx {

// ------------------------------------------------------------------------------------------ FP32 MUX2 kernel
// k_mux<K, R>: complex64 passes whose ops are all MUX2 records (and diagonal tables), in FP32.
// The workgroup's 2^K-amplitude tile sits in padded LDS as in k_pass; each thread owns one slot of
// 2^R amplitudes per register group (NT = 2^(K-R) threads).  A group's ops run on two register
// buffers in turn -- op 2i reads A and writes B, op 2i + 1 reads B and writes A -- so every result
// is computed straight into the registers it lives in (no register copies between ops), and every
// product is a two-operand FMA with the op's coefficient in a scalar register (the host stores the
// real 4x4 form of each 2x2 op with its signs folded in).
namespace mux {

template <int TB>
__device__ __forceinline__ int pr0(int p) { return ((p >> TB) << (TB + 1)) | (p & ((1 << TB) - 1)); }

// y = W x on every pair, W the signed real 4x4 form (rows y0r y0i y1r y1i, columns x0r x0i x1r x1i)
template <int R, int TB>
__device__ __forceinline__ void full(const float (&x)[2 << R], float (&y)[2 << R], const float *w) {
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int a = 2 * pr0<TB>(p), b = a + (2 << TB);
        const float x0 = x[a], x1 = x[a + 1], x2 = x[b], x3 = x[b + 1];
        y[a] = fmaf(w[0], x0, fmaf(w[1], x1, fmaf(w[2], x2, w[3] * x3)));
        y[a + 1] = fmaf(w[4], x0, fmaf(w[5], x1, fmaf(w[6], x2, w[7] * x3)));
        y[b] = fmaf(w[8], x0, fmaf(w[9], x1, fmaf(w[10], x2, w[11] * x3)));
        y[b + 1] = fmaf(w[12], x0, fmaf(w[13], x1, fmaf(w[14], x2, w[15] * x3)));
    }
}
// The same on the FP32 matrix cores.  v_mfma_f32_4x4x1_16b_f32 computes 16 independent 4 x 4
// blocks D[i][j] += A[i][k] B[k][j] with block b on lanes 4 b .. 4 b + 3: lane 4 b + i supplies
// A[i][k], lane 4 b + j supplies B[k][j] and receives D[0..3][j].  With every lane supplying its own
// pair's input component k as B and row (lane & 3) of W as A (wl[k] = W[lane & 3][k]), four steps
// leave each lane's pair (y0r, y0i, y1r, y1i) in its accumulator: FP32 products, each step one FMA
// rounding.
typedef float mf4 __attribute__((ext_vector_type(4)));
template <int R, int TB>
__device__ __forceinline__ void full_mfma(const float (&x)[2 << R], float (&y)[2 << R], const float (&wl)[4]) {
    constexpr int NP = 1 << (R - 1);
    mf4 acc[NP];
    // K step outside, pairs inside: NP independent accumulation chains in flight
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const int a = 2 * pr0<TB>(p);
        acc[p] = __builtin_amdgcn_mfma_f32_4x4x1f32(wl[0], x[a], mf4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const int a = 2 * pr0<TB>(p);
        acc[p] = __builtin_amdgcn_mfma_f32_4x4x1f32(wl[1], x[a + 1], acc[p], 0, 0, 0);
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const int b = 2 * pr0<TB>(p) + (2 << TB);
        acc[p] = __builtin_amdgcn_mfma_f32_4x4x1f32(wl[2], x[b], acc[p], 0, 0, 0);
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const int a = 2 * pr0<TB>(p), b = a + (2 << TB);
        acc[p] = __builtin_amdgcn_mfma_f32_4x4x1f32(wl[3], x[b + 1], acc[p], 0, 0, 0);
        y[a] = acc[p][0];
        y[a + 1] = acc[p][1];
        y[b] = acc[p][2];
        y[b + 1] = acc[p][3];
    }
}
// row (lane & 3) of W for the matrix-core form
__device__ __forceinline__ void lane_row(const float *w, float (&wl)[4]) {
    const uint32_t i = __lane_id();
    const bool b0 = (i & 1u) != 0, b1 = (i & 2u) != 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float lo = b0 ? w[4 + k] : w[k], hi = b0 ? w[12 + k] : w[8 + k];
        wl[k] = b1 ? hi : lo;
    }
}
// real 2x2 (w = m00, m01, m10, m11): re and im rows separately
template <int R, int TB>
__device__ __forceinline__ void real(const float (&x)[2 << R], float (&y)[2 << R], const float *w) {
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int a = 2 * pr0<TB>(p), b = a + (2 << TB);
        const float x0 = x[a], x1 = x[a + 1], x2 = x[b], x3 = x[b + 1];
        y[a] = fmaf(w[0], x0, w[1] * x2);
        y[a + 1] = fmaf(w[0], x1, w[1] * x3);
        y[b] = fmaf(w[2], x0, w[3] * x2);
        y[b + 1] = fmaf(w[2], x1, w[3] * x3);
    }
}
__device__ __forceinline__ void cmul_to(float er, float ei, float xr, float xi, float &yr, float &yi) {
    yr = fmaf(er, xr, -ei * xi);
    yi = fmaf(er, xi, ei * xr);
}
template <int R, int TB>
__device__ __forceinline__ void diag(const float (&x)[2 << R], float (&y)[2 << R], const float *w) {
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int a = 2 * pr0<TB>(p), b = a + (2 << TB);
        const float x0 = x[a], x1 = x[a + 1], x2 = x[b], x3 = x[b + 1];
        cmul_to(w[0], w[1], x0, x1, y[a], y[a + 1]);
        cmul_to(w[2], w[3], x2, x3, y[b], y[b + 1]);
    }
}
template <int R, int TB>
__device__ __forceinline__ void xswap(const float (&x)[2 << R], float (&y)[2 << R]) {
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int a = 2 * pr0<TB>(p), b = a + (2 << TB);
        const float x0 = x[a], x1 = x[a + 1], x2 = x[b], x3 = x[b + 1];
        y[a] = x2;
        y[a + 1] = x3;
        y[b] = x0;
        y[b + 1] = x1;
    }
}
// controlled ops: pair p is selected when its register bits cover cmr and the thread's other tile
// bits cover the rest of the control mask (ts)
template <int R, int TB>
__device__ __forceinline__ void cx(const float (&x)[2 << R], float (&y)[2 << R], uint32_t cmr, bool ts) {
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = pr0<TB>(p), a = 2 * r0, b = a + (2 << TB);
        const bool sel = ((uint32_t(r0) & cmr) == cmr) && ts;
        // values first: a select between two array elements would become a select of the index
        const float x0 = x[a], x1 = x[a + 1], x2 = x[b], x3 = x[b + 1];
        y[a] = sel ? x2 : x0;
        y[a + 1] = sel ? x3 : x1;
        y[b] = sel ? x0 : x2;
        y[b + 1] = sel ? x1 : x3;
    }
}
// one register control CB: the selected pairs are fixed at compile time
template <int R, int TB, int CB>
__device__ __forceinline__ void cx1(const float (&x)[2 << R], float (&y)[2 << R]) {
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = pr0<TB>(p), a = 2 * r0, b = a + (2 << TB);
        const float x0 = x[a], x1 = x[a + 1], x2 = x[b], x3 = x[b + 1];
        const bool sel = (r0 >> CB) & 1;
        y[a] = sel ? x2 : x0;
        y[a + 1] = sel ? x3 : x1;
        y[b] = sel ? x0 : x2;
        y[b + 1] = sel ? x1 : x3;
    }
}
template <int R, int TB, int CB>
__device__ __forceinline__ void cphase1(const float (&x)[2 << R], float (&y)[2 << R], const float *w) {
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = pr0<TB>(p), a = 2 * r0, b = a + (2 << TB);
        const float x0 = x[a], x1 = x[a + 1], x2 = x[b], x3 = x[b + 1];
        y[a] = x0;
        y[a + 1] = x1;
        if ((r0 >> CB) & 1) {
            cmul_to(w[0], w[1], x2, x3, y[b], y[b + 1]);
        } else {
            y[b] = x2;
            y[b + 1] = x3;
        }
    }
}
// controls only outside the register bits: one lane mask (ts) for every pair
template <int R, int TB>
__device__ __forceinline__ void cxt(const float (&x)[2 << R], float (&y)[2 << R], bool ts) {
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int a = 2 * pr0<TB>(p), b = a + (2 << TB);
        const float x0 = x[a], x1 = x[a + 1], x2 = x[b], x3 = x[b + 1];
        y[a] = ts ? x2 : x0;
        y[a + 1] = ts ? x3 : x1;
        y[b] = ts ? x0 : x2;
        y[b + 1] = ts ? x1 : x3;
    }
}
template <int R, int TB>
__device__ __forceinline__ void cphaset(const float (&x)[2 << R], float (&y)[2 << R], const float *w, bool ts) {
    // the selected lanes multiply by e, the others by 1
    const float er = ts ? w[0] : 1.f, ei = ts ? w[1] : 0.f;
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int a = 2 * pr0<TB>(p), b = a + (2 << TB);
        const float x0 = x[a], x1 = x[a + 1], x2 = x[b], x3 = x[b + 1];
        y[a] = x0;
        y[a + 1] = x1;
        cmul_to(er, ei, x2, x3, y[b], y[b + 1]);
    }
}
template <int R, int TB>
__device__ __forceinline__ void cphase(const float (&x)[2 << R], float (&y)[2 << R], const float *w, uint32_t cmr,
                                       bool ts) {
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = pr0<TB>(p), a = 2 * r0, b = a + (2 << TB);
        const bool sel = ((uint32_t(r0) & cmr) == cmr) && ts;
        const float x0 = x[a], x1 = x[a + 1], x2 = x[b], x3 = x[b + 1];
        float pr, pi;
        cmul_to(w[0], w[1], x2, x3, pr, pi);
        y[a] = x0;
        y[a + 1] = x1;
        y[b] = sel ? pr : x2;
        y[b + 1] = sel ? pi : x3;
    }
}
// [y0, y1] = M [x0, x1], M complex (m00 m01 m10 m11 as re, im)
__device__ __forceinline__ void mat2_to(const float *m, float x0r, float x0i, float x1r, float x1i, float &y0r,
                                        float &y0i, float &y1r, float &y1i) {
    y0r = fmaf(m[0], x0r, fmaf(-m[1], x0i, fmaf(m[2], x1r, -m[3] * x1i)));
    y0i = fmaf(m[0], x0i, fmaf(m[1], x0r, fmaf(m[2], x1i, m[3] * x1r)));
    y1r = fmaf(m[4], x0r, fmaf(-m[5], x0i, fmaf(m[6], x1r, -m[7] * x1i)));
    y1i = fmaf(m[4], x0i, fmaf(m[5], x0r, fmaf(m[6], x1i, m[7] * x1r)));
}
template <int R, int TB>
__device__ __forceinline__ void cgen(const float (&x)[2 << R], float (&y)[2 << R], const float *m0, const float *m1,
                                     uint32_t cmr, bool ts) {
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = pr0<TB>(p), a = 2 * r0, b = a + (2 << TB);
        const bool sel = ((uint32_t(r0) & cmr) == cmr) && ts;
        const float x0 = x[a], x1 = x[a + 1], x2 = x[b], x3 = x[b + 1];
        float u0, u1, u2, u3, v0, v1, v2, v3;
        mat2_to(m0, x0, x1, x2, x3, u0, u1, u2, u3);
        mat2_to(m1, x0, x1, x2, x3, v0, v1, v2, v3);
        y[a] = sel ? v0 : u0;
        y[a + 1] = sel ? v1 : u1;
        y[b] = sel ? v2 : u2;
        y[b + 1] = sel ? v3 : u3;
    }
}
// diagonal table on the registers: register r of slot s times entry r NSLOT + s
template <int R, uint32_t NSLOT>
__device__ __forceinline__ void table(const float (&x)[2 << R], float (&y)[2 << R], const float *__restrict__ tab,
                                      uint32_t s) {
    const float2 *d = reinterpret_cast<const float2 *>(tab) + s;
#pragma unroll
    for (int r0 = 0; r0 < (1 << R); r0 += 8) {
        float2 t[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) t[r] = d[(r0 + r) * NSLOT];
#pragma unroll
        for (int r = 0; r < 8; ++r) cmul_to(t[r].x, t[r].y, x[2 * (r0 + r)], x[2 * (r0 + r) + 1], y[2 * (r0 + r)], y[2 * (r0 + r) + 1]);
    }
}

// One op from its 96-byte record (kind = (M_* - M_FULL) * 8 + target register bit, cmr, cmt, aux,
// then 16 coefficients), already in scalar registers.
struct Rec {
    uint4 h0, h1, c0, c1, c2, c3;
};
__device__ __forceinline__ Rec load_rec(const OpDev<float> *op) {
    const uint4 *p = reinterpret_cast<const uint4 *>(op);
    return Rec{p[0], p[1], p[2], p[3], p[4], p[5]};
}

// single register control CB of op kind CX1 / CPHASE1 (CB != TB, CB < R: the host guarantees it)
template <int R, int TB, bool PHASE>
__device__ __forceinline__ void ctl1(int cb, const float *w, const float (&x)[2 << R], float (&y)[2 << R]) {
#define SVX_C1(C)                                                                                  \
    if constexpr (C != TB && C < R) {                                                              \
        if (cb == C) {                                                                             \
            if constexpr (PHASE) cphase1<R, TB, C>(x, y, w); else cx1<R, TB, C>(x, y);             \
            return;                                                                                \
        }                                                                                          \
    }
    SVX_C1(0) SVX_C1(1) SVX_C1(2) SVX_C1(3) SVX_C1(4)
#undef SVX_C1
    // not reached (cb is one of the arms above); a valid arm keeps every path writing y
    if constexpr (PHASE) cphase1<R, TB, (TB == R - 1 ? R - 2 : R - 1)>(x, y, w);
    else cx1<R, TB, (TB == R - 1 ? R - 2 : R - 1)>(x, y);
}

template <int R, int TB>
__device__ __forceinline__ void kind_tb(int kind, const float *w, uint32_t cmr, const float (&x)[2 << R],
                                        float (&y)[2 << R], bool ts) {
    switch (kind) {
    case M_CX1 - M_FULL: ctl1<R, TB, false>((int)cmr, w, x, y); break;
    case M_CXT - M_FULL: cxt<R, TB>(x, y, ts); break;
    case M_CPHASE1 - M_FULL: ctl1<R, TB, true>((int)cmr, w, x, y); break;
    case M_CPHASET - M_FULL: cphaset<R, TB>(x, y, w, ts); break;
    case M_FULL - M_FULL: full<R, TB>(x, y, w); break;
    case M_REAL - M_FULL: real<R, TB>(x, y, w); break;
    case M_DIAG - M_FULL: diag<R, TB>(x, y, w); break;
    case M_X - M_FULL: xswap<R, TB>(x, y); break;
    case M_CX - M_FULL: cx<R, TB>(x, y, cmr, ts); break;
    case M_CPHASE - M_FULL: cphase<R, TB>(x, y, w, cmr, ts); break;
    case M_FULLM - M_FULL: {
        float wl[4];
        lane_row(w, wl);
        full_mfma<R, TB>(x, y, wl);
        break;
    }
    default: cgen<R, TB>(x, y, w, w + 8, cmr, ts); break;
    }
}

template <int R, uint32_t NSLOT>
__device__ __forceinline__ void apply(const Rec &rc, const float (&x)[2 << R], float (&y)[2 << R], uint32_t base,
                                      const float *__restrict__ wd, uint32_t s) {
    const int code = (int)rc.h0.x;
    const uint32_t cmr = rc.h0.z, cmt = rc.h0.w;
    if (code == (M_TABLE - M_FULL) * 8) {
        table<R, NSLOT>(x, y, wd + rc.h1.z, s);
        return;
    }
    float w[16] = {__uint_as_float(rc.c0.x), __uint_as_float(rc.c0.y), __uint_as_float(rc.c0.z), __uint_as_float(rc.c0.w),
                   __uint_as_float(rc.c1.x), __uint_as_float(rc.c1.y), __uint_as_float(rc.c1.z), __uint_as_float(rc.c1.w),
                   __uint_as_float(rc.c2.x), __uint_as_float(rc.c2.y), __uint_as_float(rc.c2.z), __uint_as_float(rc.c2.w),
                   __uint_as_float(rc.c3.x), __uint_as_float(rc.c3.y), __uint_as_float(rc.c3.z), __uint_as_float(rc.c3.w)};
    const bool ts = (base & cmt) == cmt;
    const int kind = code >> 3, tr = code & 7;
    // every arm writes all of y (an arm that left y alone would keep both buffers live); the host
    // only emits target register bits 0 .. R-1
    if constexpr (R == 5) {
        switch (tr) {
        case 0: kind_tb<R, 0>(kind, w, cmr, x, y, ts); break;
        case 1: kind_tb<R, 1>(kind, w, cmr, x, y, ts); break;
        case 2: kind_tb<R, 2>(kind, w, cmr, x, y, ts); break;
        case 3: kind_tb<R, 3>(kind, w, cmr, x, y, ts); break;
        default: kind_tb<R, 4>(kind, w, cmr, x, y, ts); break;
        }
    } else {
        switch (tr) {
        case 0: kind_tb<R, 0>(kind, w, cmr, x, y, ts); break;
        case 1: kind_tb<R, 1>(kind, w, cmr, x, y, ts); break;
        case 2: kind_tb<R, 2>(kind, w, cmr, x, y, ts); break;
        default: kind_tb<R, 3>(kind, w, cmr, x, y, ts); break;
        }
    }
}

}  // namespace mux

// One workgroup tile of a MUX2 pass (k_mux and its batched form): NT = 2^(K-R) threads, one slot each.
// DN: the pass also holds FP32 dense units (group types 1 and 5), applied by all NT / 64 wavefronts
// on the padded LDS tile between the register groups.
// zero: the tile is the whole state and starts as |0...0> (written to LDS, nothing read from HBM).
template <int K, int R, bool DN>
__device__ __forceinline__ void mux_tile(C2<float> *__restrict__ st, const PassDev *__restrict__ pd,
                                         const GroupDev *__restrict__ gd, const OpDev<float> *__restrict__ od,
                                         const float *__restrict__ wd, uint64_t tile_id, unsigned char *smem,
                                         bool zero = false) {
    constexpr int NT = 1 << (K - R);
    constexpr uint32_t NE = 1u << K;
    constexpr uint32_t NEP = NE + (NE >> 5);
    constexpr int NR = 1 << R;
    C2<float> *tile = reinterpret_cast<C2<float> *>(smem);
    uint64_t *sbo = reinterpret_cast<uint64_t *>(smem + NEP * sizeof(C2<float>));
    const int k = pd->k, m = pd->m;
    const int sub = K - k;
    uint64_t *tb = sbo + (1u << m);
    const uint32_t tid = threadIdx.x;
    for (uint32_t q = tid; q < (1u << m); q += NT) {
        uint64_t o = 0;
        for (int i = 0; i < m; ++i)
            if ((q >> i) & 1u) o |= uint64_t(1) << pd->h[i];
        sbo[q] = o;
    }
    for (uint32_t u = tid; u < (1u << sub); u += NT) {
        uint64_t x = ((tile_id << sub) | u) << (k - m);
        for (int i = 0; i < m; ++i) x = ins0_64(x, pd->hs[i]);
        tb[u] = x;
    }
    __syncthreads();
    const int lowb = k - m;
    const int ng = pd->ngroups, g0 = pd->group0;
    // a three-product dense unit's A operands are loaded one group ahead (before the barrier that
    // precedes the unit, or with the tile), so their latency overlaps the barrier or the tile load
    Gauss4W gw;
    int gw_g = -1;   // the group whose operands gw holds
    auto prefetch = [&](int g) {
        if constexpr (DN) {
            if (g < ng && gd[g0 + g].type == 5) {
                gauss4_weights(wd + gd[g0 + g].w_off, tid, gw);
                gw_g = g;
            }
        }
    };
    prefetch(0);
    if (zero) {
        for (uint32_t e = tid; e < NE; e += NT) tile[lpad(e)] = C2<float>{e == 0 ? 1.0f : 0.0f, 0.0f};
    } else {
        move_tile<float, K, true, NT>(st, tile, sbo, tb, k, lowb, tid);
    }
    __syncthreads();
    for (int g = 0; g < ng; ++g) {
        const GroupDev &G = gd[g0 + g];
        if (G.type == 6) {
            diag_table<K, NT>(tile, wd + G.w_off, k, tid);
            prefetch(g + 1);
            __syncthreads();
            continue;
        }
        if constexpr (DN) {
            if (G.type != 0) {
                if (G.type == 5) {
                    if (gw_g != g) prefetch(g);
                    dense_gauss4<K, NT / 64, false>(smem, G, gw, tid);
                } else {
                    dense_group_f32<K, NT / 64, false>(smem, G, wd, tid);
                }
                prefetch(g + 1);
                __syncthreads();
                continue;
            }
        }
        uint32_t poff[NR];   // padded LDS offset of register r (disjoint bits: pad(a|b) = pad(a)+pad(b))
        poff[0] = 0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const uint32_t bq = lpad(1u << G.b[q]);
#pragma unroll
            for (int r = (1 << q); r < (2 << q); ++r) poff[r] = poff[r - (1 << q)] + bq;
        }
        uint32_t base = tid;
#pragma unroll
        for (int q = 0; q < R; ++q) base = ins0_32(base, G.bs[q]);
        const uint32_t pb = lpad(base);
        float A[2 * NR], B[2 * NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const C2<float> v = tile[pb + poff[r]];
            A[2 * r] = v.x;
            A[2 * r + 1] = v.y;
        }
        const int nops = G.nops;
        const OpDev<float> *ops = od + G.op0;
        int o = 0;
#pragma unroll 1
        for (; o + 1 < nops; o += 2) {
            mux::apply<R, NE / NR>(mux::load_rec(ops + o), A, B, base, wd, tid);
            mux::apply<R, NE / NR>(mux::load_rec(ops + o + 1), B, A, base, wd, tid);
        }
        if (o < nops) {
            mux::apply<R, NE / NR>(mux::load_rec(ops + o), A, B, base, wd, tid);
#pragma unroll
            for (int r = 0; r < NR; ++r) tile[pb + poff[r]] = C2<float>{B[2 * r], B[2 * r + 1]};
        } else {
#pragma unroll
            for (int r = 0; r < NR; ++r) tile[pb + poff[r]] = C2<float>{A[2 * r], A[2 * r + 1]};
        }
        prefetch(g + 1);
        __syncthreads();
    }
    move_tile<float, K, false, NT>(st, tile, sbo, tb, k, lowb, tid);
}

// Occupancy: the second __launch_bounds__ argument is the minimum number of wavefronts per SIMD.
// R = 4 (K <= 13): four per SIMD, so at most 128 VGPRs -- at K = 13 two eight-wavefront
// workgroups per CU (LDS holds two 64 KiB tiles), at smaller K more and smaller ones.  K = 14
// (R = 5): one workgroup per CU (one 128 KiB tile), two wavefronts per SIMD.  The batched form
// reads its descriptors through the constant address space and fits the same bound; the queued
// form keeps the compiler's choice (its descriptors are read with vector loads, and held to 128
// VGPRs it spills).
template <int K, int R, bool DN>
__global__ __launch_bounds__(1 << (K - R), R == 4 ? 4 : 2) void k_mux(C2<float> *__restrict__ st, const PassDev *__restrict__ pd,
                                                      const GroupDev *__restrict__ gd,
                                                      const OpDev<float> *__restrict__ od, const float *__restrict__ wd) {
    extern __shared__ __align__(16) unsigned char smem[];
    mux_tile<K, R, DN>(st, pd, gd, od, wd, tile_order(blockIdx.x, pd), smem);
}

template <int K, int R, bool DN>
__global__ __launch_bounds__(1 << (K - R), R == 4 ? 4 : 2) void k_mux_batch(const BatchItem *__restrict__ items, int n_items) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint64_t b = blockIdx.x;
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {   // last item with tile0 <= b
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].tile0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const BatchItem &it = items[__builtin_amdgcn_readfirstlane(lo)];
    // the item's pointers are workgroup-uniform: keep them in scalar registers (the op records are
    // read with scalar loads)
    auto uni = [](auto *ptr) {
        const uint64_t v = reinterpret_cast<uint64_t>(ptr);
        const uint64_t u = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v) |
                           ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32);
        return reinterpret_cast<decltype(ptr)>(u);
    };
    // the descriptors are read-only for the launch: read them through the constant address space,
    // so that the compiler keeps uniform descriptor words in scalar registers as it does for k_mux's
    // kernel arguments
    auto cst = [](auto *ptr) {
        using T = std::remove_cv_t<std::remove_pointer_t<decltype(ptr)>>;
        const __attribute__((address_space(4))) T *c =
            (const __attribute__((address_space(4))) T *)reinterpret_cast<uintptr_t>(ptr);
        return (const T *)c;
    };
    const PassDev *pd = cst(uni(it.pd));
    const uint64_t t0 = uni(reinterpret_cast<const char *>(it.tile0)) - static_cast<const char *>(nullptr);
    mux_tile<K, R, DN>(reinterpret_cast<C2<float> *>(uni(it.st)), pd, cst(uni(it.gd)),
                   cst(reinterpret_cast<const OpDev<float> *>(uni(it.od))), cst(uni(it.wd)), tile_order(b - t0, pd), smem);
}

// Persistent queue of whole-state circuits (state = one 2^K tile, one pass each).  The host
// publishes jobs in host-visible memory (QueueJob records, then QueueCtl::tail with a release
// store); each workgroup claims the next published job with a compare-and-swap on a device counter
// (never an unpublished one), runs it with the same device code as k_mux, makes its stores visible
// at device scope, and sets the job's done word (ticket + 1) with a system-scope release store.  A
// workgroup leaves when the host sets QueueCtl::stop, when no job has arrived for idle_ticks of
// the 100 MHz real-time counter, or after 2^24 idle rounds, so every launch drains by itself; the
// host relaunches on demand.
template <int K, int R>
__global__ __launch_bounds__(1 << (K - R)) void k_mux_queue(const QueueJob *jobs, QueueCtl *ctl, uint32_t *next,
                                                            uint32_t cap, uint64_t idle_ticks) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint64_t s_job[7];
    constexpr uint32_t kNone = 0xffffffffu;
    while (true) {
        if (threadIdx.x == 0) {
            uint32_t claimed = kNone;
            uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            // each idle round reads host memory and sleeps; the round cap bounds the wait even if the
            // clock does not advance
            for (uint32_t round = 0; round < (1u << 24); ++round) {
                const uint32_t tail = __hip_atomic_load(&ctl->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
                uint32_t cur = __hip_atomic_load(next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (cur < tail) {
                    if (__hip_atomic_compare_exchange_strong(next, &cur, cur + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) {
                        claimed = cur;
                        break;
                    }
                    continue;
                }
                if (__hip_atomic_load(&ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) break;
                if (__builtin_amdgcn_s_memrealtime() - t0 > idle_ticks) break;
                __builtin_amdgcn_s_sleep(8);
            }
            s_job[0] = claimed;
            if (claimed != kNone) {
                const QueueJob &jb = jobs[claimed % cap];
                s_job[1] = __hip_atomic_load(&jb.st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                s_job[2] = __hip_atomic_load(&jb.pd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                s_job[3] = __hip_atomic_load(&jb.gd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                s_job[4] = __hip_atomic_load(&jb.od, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                s_job[5] = __hip_atomic_load(&jb.wd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                s_job[6] = __hip_atomic_load(&jb.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
        __syncthreads();
        const uint32_t j = (uint32_t)s_job[0];
        if (j == kNone) return;   // the whole workgroup leaves together
        const uint64_t flags = s_job[6];
        auto *st = reinterpret_cast<C2<float> *>(s_job[1]);
        const auto *pd = reinterpret_cast<const PassDev *>(s_job[2]);
        const auto *gd = reinterpret_cast<const GroupDev *>(s_job[3]);
        const auto *od = reinterpret_cast<const OpDev<float> *>(s_job[4]);
        const auto *wd = reinterpret_cast<const float *>(s_job[5]);
        mux_tile<K, R, true>(st, pd, gd, od, wd, 0, smem, (flags & 1u) != 0);
        __threadfence();   // this thread's state stores reach device scope
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(&ctl->done[j % cap], j + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        __syncthreads();   // s_job is rewritten by the next claim
    }
}

hipError_t launch_mux_queue(int K, const QueueJob *jobs, QueueCtl *ctl, uint32_t *next, uint32_t cap,
                            uint64_t idle_ticks, uint32_t grid, size_t lds, hipStream_t s) {
    switch (K) {
    case 10: hipLaunchKernelGGL((k_mux_queue<10, 4>), dim3(grid), dim3(1 << 6), lds, s, jobs, ctl, next, cap, idle_ticks); break;
    case 11: hipLaunchKernelGGL((k_mux_queue<11, 4>), dim3(grid), dim3(1 << 7), lds, s, jobs, ctl, next, cap, idle_ticks); break;
    case 12: hipLaunchKernelGGL((k_mux_queue<12, 4>), dim3(grid), dim3(1 << 8), lds, s, jobs, ctl, next, cap, idle_ticks); break;
    case 13: hipLaunchKernelGGL((k_mux_queue<13, kMuxR13>), dim3(grid), dim3(1 << (13 - kMuxR13)), lds, s, jobs, ctl, next, cap, idle_ticks); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// register bits of the k_mux instance for a workgroup tile of 2^K amplitudes
int mux_register_bits(int K) { return K >= 14 ? 5 : (K == 13 ? kMuxR13 : 4); }

#define SVX_MUX_CASES(DN)                                                                                   \
    switch (K) {                                                                                            \
    case 10: hipLaunchKernelGGL((k_mux<10, 4, DN>), dim3(grid), dim3(1 << 6), lds, s, p, pd, gd, od, wd); break; \
    case 11: hipLaunchKernelGGL((k_mux<11, 4, DN>), dim3(grid), dim3(1 << 7), lds, s, p, pd, gd, od, wd); break; \
    case 12: hipLaunchKernelGGL((k_mux<12, 4, DN>), dim3(grid), dim3(1 << 8), lds, s, p, pd, gd, od, wd); break; \
    case 13: hipLaunchKernelGGL((k_mux<13, kMuxR13, DN>), dim3(grid), dim3(1 << (13 - kMuxR13)), lds, s, p, pd, gd, od, wd); break; \
    case 14: hipLaunchKernelGGL((k_mux<14, 5, DN>), dim3(grid), dim3(1 << 9), lds, s, p, pd, gd, od, wd); break; \
    default: return hipErrorInvalidValue;                                                                   \
    }

hipError_t launch_mux(int K, bool dense, void *st, const PassDev *pd, const GroupDev *gd, const OpDev<float> *od,
                      const float *wd, uint32_t grid, size_t lds, hipStream_t s) {
    C2<float> *p = reinterpret_cast<C2<float> *>(st);
    if (dense) {
        SVX_MUX_CASES(true)
    } else {
        SVX_MUX_CASES(false)
    }
    return hipGetLastError();
}
#undef SVX_MUX_CASES

#define SVX_MUXB_CASES(DN)                                                                                   \
    switch (K) {                                                                                             \
    case 10: hipLaunchKernelGGL((k_mux_batch<10, 4, DN>), dim3(grid), dim3(1 << 6), lds, s, items, n_items); break; \
    case 11: hipLaunchKernelGGL((k_mux_batch<11, 4, DN>), dim3(grid), dim3(1 << 7), lds, s, items, n_items); break; \
    case 12: hipLaunchKernelGGL((k_mux_batch<12, 4, DN>), dim3(grid), dim3(1 << 8), lds, s, items, n_items); break; \
    case 13: hipLaunchKernelGGL((k_mux_batch<13, kMuxR13, DN>), dim3(grid), dim3(1 << (13 - kMuxR13)), lds, s, items, n_items); break; \
    case 14: hipLaunchKernelGGL((k_mux_batch<14, 5, DN>), dim3(grid), dim3(1 << 9), lds, s, items, n_items); break; \
    default: return hipErrorInvalidValue;                                                                    \
    }

hipError_t launch_mux_batch(int K, bool dense, const BatchItem *items, int n_items, uint32_t grid, size_t lds,
                            hipStream_t s) {
    if (dense) {
        SVX_MUXB_CASES(true)
    } else {
        SVX_MUXB_CASES(false)
    }
    return hipGetLastError();
}
#undef SVX_MUXB_CASES

hipError_t set_mux_attributes() {
    const int big = 160 * 1024;
    hipError_t e = hipSuccess;
#define SVX_ATTR(fn)                                                                              \
    if (e == hipSuccess) e = hipFuncSetAttribute(reinterpret_cast<const void *>(fn),               \
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, big);
#define SVX_ATTR_DN(DN)                                                                                      \
    SVX_ATTR((k_mux<10, 4, DN>)) SVX_ATTR((k_mux<11, 4, DN>)) SVX_ATTR((k_mux<12, 4, DN>))                  \
    SVX_ATTR((k_mux<13, kMuxR13, DN>)) SVX_ATTR((k_mux<14, 5, DN>)) SVX_ATTR((k_mux_batch<10, 4, DN>))        \
    SVX_ATTR((k_mux_batch<11, 4, DN>)) SVX_ATTR((k_mux_batch<12, 4, DN>)) SVX_ATTR((k_mux_batch<13, kMuxR13, DN>)) \
    SVX_ATTR((k_mux_batch<14, 5, DN>))
    SVX_ATTR_DN(false)
    SVX_ATTR_DN(true)
    // the queue kernel also holds a little static LDS (its job record)
#define SVX_ATTRQ(fn)                                                                             \
    if (e == hipSuccess) e = hipFuncSetAttribute(reinterpret_cast<const void *>(fn),               \
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, big - 1024);
    SVX_ATTRQ((k_mux_queue<10, 4>)) SVX_ATTRQ((k_mux_queue<11, 4>)) SVX_ATTRQ((k_mux_queue<12, 4>))
    SVX_ATTRQ((k_mux_queue<13, kMuxR13>))
#undef SVX_ATTRQ
#undef SVX_ATTR_DN
#undef SVX_ATTR
    return e;
}

}  // namespace svx
