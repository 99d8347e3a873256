// Device kernels of libsvx (gfx950).
//
// Pass kernel (k_pass<T, K>): one workgroup owns a 2^K-amplitude tile (one or more QSV2 tiles) in
// LDS.  It gathers the tile from HBM with 16-byte loads (2^m sub-blocks of 2^(k-m) contiguous
// amplitudes, the QSV2 tile of kernels/sv/sv_engine.cpp:86-118 / host/sv/model.py:25-40), applies
// the pass's MUX2 ops group by group -- each group is a run of consecutive ops whose targets lie in
// 5 tile bits; every thread pulls the 2^5 amplitudes of one slot into registers, applies all of the
// group's ops there (targets are compile-time register indices, controls are a uniform register
// mask plus a per-thread predicate), and writes them back -- and stores the tile once.
// MUX2 semantics (host/sv/isa.py:17-21): pair (i0, i0 | 2^t), bit t of i0 clear, uses M1 iff
// cmask != 0 and (i0 & cmask) == cmask, else M0.
#include "svx_device.h"

namespace svx {

// [a, b] <- M [a, b], M = (m00 m01 m10 m11) as (re, im)
template <typename T>
__device__ __forceinline__ void mat2(const T *m, C2<T> &a, C2<T> &b) {
    const C2<T> x0 = a, x1 = b;
    a.x = fmaT(m[0], x0.x, fmaT(-m[1], x0.y, fmaT(m[2], x1.x, -m[3] * x1.y)));
    a.y = fmaT(m[0], x0.y, fmaT(m[1], x0.x, fmaT(m[2], x1.y, m[3] * x1.x)));
    b.x = fmaT(m[4], x0.x, fmaT(-m[5], x0.y, fmaT(m[6], x1.x, -m[7] * x1.y)));
    b.y = fmaT(m[4], x0.y, fmaT(m[5], x0.x, fmaT(m[6], x1.y, m[7] * x1.x)));
}

template <typename T>
__device__ __forceinline__ C2<T> cmul(T er, T ei, C2<T> v) {
    C2<T> r;
    r.x = fmaT(er, v.x, -ei * v.y);
    r.y = fmaT(er, v.y, ei * v.x);
    return r;
}

// ------------------------------------------------------------------------------------------ ops
template <typename T, int R, int TB>
__device__ __forceinline__ void op_full(C2<T> (&a)[1 << R], const T *m0) {
    constexpr int TBIT = 1 << TB;
    T mm[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) mm[e] = m0[e];
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = ((p >> TB) << (TB + 1)) | (p & (TBIT - 1));
        mat2(mm, a[r0], a[r0 | TBIT]);
    }
}
template <typename T, int R, int TB>
__device__ __forceinline__ void op_diag(C2<T> (&a)[1 << R], const T *m0) {
    constexpr int TBIT = 1 << TB;
    const T d0r = m0[0], d0i = m0[1], d1r = m0[6], d1i = m0[7];
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = ((p >> TB) << (TB + 1)) | (p & (TBIT - 1));
        a[r0] = cmul(d0r, d0i, a[r0]);
        a[r0 | TBIT] = cmul(d1r, d1i, a[r0 | TBIT]);
    }
}
template <typename T, int R, int TB>
__device__ __forceinline__ void op_xswap(C2<T> (&a)[1 << R]) {
    constexpr int TBIT = 1 << TB;
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = ((p >> TB) << (TB + 1)) | (p & (TBIT - 1));
        const C2<T> t0 = a[r0];
        a[r0] = a[r0 | TBIT];
        a[r0 | TBIT] = t0;
    }
}
template <typename T, int R, int TB>
__device__ __forceinline__ void op_cfull(C2<T> (&a)[1 << R], const T *m1, uint32_t cmr, bool ts) {
    constexpr int TBIT = 1 << TB;
    T me[8];
    me[0] = ts ? m1[0] : T(1);
    me[1] = ts ? m1[1] : T(0);
    me[2] = ts ? m1[2] : T(0);
    me[3] = ts ? m1[3] : T(0);
    me[4] = ts ? m1[4] : T(0);
    me[5] = ts ? m1[5] : T(0);
    me[6] = ts ? m1[6] : T(1);
    me[7] = ts ? m1[7] : T(0);
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = ((p >> TB) << (TB + 1)) | (p & (TBIT - 1));
        if ((uint32_t(r0) & cmr) == cmr) mat2(me, a[r0], a[r0 | TBIT]);
    }
}
template <typename T, int R, int TB>
__device__ __forceinline__ void op_cdiag(C2<T> (&a)[1 << R], const T *m1, uint32_t cmr, bool ts) {
    constexpr int TBIT = 1 << TB;
    const T d0r = ts ? m1[0] : T(1), d0i = ts ? m1[1] : T(0);
    const T d1r = ts ? m1[6] : T(1), d1i = ts ? m1[7] : T(0);
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = ((p >> TB) << (TB + 1)) | (p & (TBIT - 1));
        if ((uint32_t(r0) & cmr) == cmr) {
            a[r0] = cmul(d0r, d0i, a[r0]);
            a[r0 | TBIT] = cmul(d1r, d1i, a[r0 | TBIT]);
        }
    }
}
template <typename T, int R, int TB>
__device__ __forceinline__ void op_cphase(C2<T> (&a)[1 << R], const T *m1, uint32_t cmr, bool ts) {
    constexpr int TBIT = 1 << TB;
    const T er = ts ? m1[6] : T(1), ei = ts ? m1[7] : T(0);
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = ((p >> TB) << (TB + 1)) | (p & (TBIT - 1));
        if ((uint32_t(r0) & cmr) == cmr) a[r0 | TBIT] = cmul(er, ei, a[r0 | TBIT]);
    }
}
template <typename T, int R, int TB>
__device__ __forceinline__ void op_cxswap(C2<T> (&a)[1 << R], uint32_t cmr, bool ts) {
    constexpr int TBIT = 1 << TB;
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = ((p >> TB) << (TB + 1)) | (p & (TBIT - 1));
        if ((uint32_t(r0) & cmr) == cmr) {
            const C2<T> t0 = a[r0], t1 = a[r0 | TBIT];
            a[r0].x = ts ? t1.x : t0.x;
            a[r0].y = ts ? t1.y : t0.y;
            a[r0 | TBIT].x = ts ? t0.x : t1.x;
            a[r0 | TBIT].y = ts ? t0.y : t1.y;
        }
    }
}
template <typename T, int R, int TB>
__device__ __forceinline__ void op_mux(C2<T> (&a)[1 << R], const T *m0p, const T *m1, uint32_t cmr, bool ts) {
    constexpr int TBIT = 1 << TB;
    T m0[8], ms[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        m0[e] = m0p[e];
        ms[e] = ts ? m1[e] : m0p[e];
    }
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = ((p >> TB) << (TB + 1)) | (p & (TBIT - 1));
        if ((uint32_t(r0) & cmr) == cmr)
            mat2(ms, a[r0], a[r0 | TBIT]);
        else
            mat2(m0, a[r0], a[r0 | TBIT]);
    }
}

// controlled ops whose single control is register bit CB (no per-thread control): only the
// pairs with bit CB set are touched, chosen at compile time
template <typename T, int R, int TB, int CB>
__device__ __forceinline__ void op_cxswap1(C2<T> (&a)[1 << R]) {
    constexpr int TBIT = 1 << TB;
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = ((p >> TB) << (TB + 1)) | (p & (TBIT - 1));
        if (r0 & (1 << CB)) {
            const C2<T> t0 = a[r0];
            a[r0] = a[r0 | TBIT];
            a[r0 | TBIT] = t0;
        }
    }
}
template <typename T, int R, int TB, int CB>
__device__ __forceinline__ void op_cphase1(C2<T> (&a)[1 << R], const T *m1) {
    constexpr int TBIT = 1 << TB;
    const T er = m1[6], ei = m1[7];
#pragma unroll
    for (int p = 0; p < (1 << (R - 1)); ++p) {
        const int r0 = ((p >> TB) << (TB + 1)) | (p & (TBIT - 1));
        if (r0 & (1 << CB)) a[r0 | TBIT] = cmul(er, ei, a[r0 | TBIT]);
    }
}
template <typename T, int R, int TB>
__device__ __forceinline__ void op_ctl1(C2<T> (&a)[1 << R], const OpDev<T> &op, bool swap) {
    // control register bit = index of the single set bit of cmr
    const int cb = __builtin_ctz(op.cmr | (1u << 31));
#define SVX_CB(CBV)                                                                        \
    case CBV:                                                                              \
        if constexpr (CBV != TB && CBV < R) {                                              \
            if (swap) op_cxswap1<T, R, TB, CBV>(a); else op_cphase1<T, R, TB, CBV>(a, op.m1); \
        }                                                                                  \
        break;
    switch (cb) { SVX_CB(0) SVX_CB(1) SVX_CB(2) SVX_CB(3) SVX_CB(4) default: break; }
#undef SVX_CB
}

// one switch over (kind, target register bit): code = kind * 8 + tr, set by the host
template <typename T, int R, int TB>
__device__ __forceinline__ void apply_kind(C2<T> (&a)[1 << R], const OpDev<T> &op, bool ts, int kind) {
    switch (kind) {
    case K_FULL: op_full<T, R, TB>(a, op.m0); break;
    case K_DIAG: op_diag<T, R, TB>(a, op.m0); break;
    case K_XSWAP: op_xswap<T, R, TB>(a); break;
    case K_CFULL: op_cfull<T, R, TB>(a, op.m1, op.cmr, ts); break;
    case K_CDIAG: op_cdiag<T, R, TB>(a, op.m1, op.cmr, ts); break;
    case K_CPHASE: op_cphase<T, R, TB>(a, op.m1, op.cmr, ts); break;
    case K_CXSWAP: op_cxswap<T, R, TB>(a, op.cmr, ts); break;
    case K_CXSWAP1: op_ctl1<T, R, TB>(a, op, true); break;
    case K_CPHASE1: op_ctl1<T, R, TB>(a, op, false); break;
    default: op_mux<T, R, TB>(a, op.m0, op.m1, op.cmr, ts); break;
    }
}

template <typename T, int R>
__device__ __forceinline__ void apply_op(C2<T> (&a)[1 << R], const OpDev<T> &op, bool ts) {
    static_assert(R == 4 || R == 5, "register groups hold 4 or 5 tile bits");
    const int kind = op.kind;
    switch (op.tr) {
    case 0: apply_kind<T, R, 0>(a, op, ts, kind); break;
    case 1: apply_kind<T, R, 1>(a, op, ts, kind); break;
    case 2: apply_kind<T, R, 2>(a, op, ts, kind); break;
    case 3: apply_kind<T, R, 3>(a, op, ts, kind); break;
    default:
        if constexpr (R == 5) apply_kind<T, R, 4>(a, op, ts, kind);
        break;
    }
}

// Workgroup maximum of mb (kNT threads, red: kNT / 64 words of LDS) -> shift that brings the tile's
// largest component below 2^target (0 for an all-zero tile).
template <int NT = kNT>
__device__ __forceinline__ int tile_shift(uint32_t mb, uint32_t *red, uint32_t tid, int target) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mb = max(mb, (uint32_t)__shfl_xor((int)mb, off));
    if ((tid & 63u) == 0) red[tid >> 6] = mb;
    __syncthreads();
    uint32_t m = red[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) m = max(m, red[w]);
    if (m == 0) return 0;
    const int efield = (int)(m >> 23);
    const int ex = efield == 0 ? -125 : efield - 126;   // largest component < 2^ex
    return target - ex;
}

// Complex64 tile load for passes with pre-scaled dense groups: the tile lands in LDS multiplied by
// 2^shift so that its largest component is below 2^target; returns the shift.  With K <= 13 and
// at least one contiguous low bit every element of the thread is loaded into registers first
// (all loads in flight at once), otherwise the tile is loaded and then rescaled in LDS.
template <int K, int NT = kNT>
__device__ __forceinline__ int load_tile_scaled(C2<float> *__restrict__ st, C2<float> *tile, const uint64_t *sbo,
                                                const uint64_t *tb, int k, int lowb, uint32_t tid, uint32_t *red,
                                                int target) {
    constexpr uint32_t NE = 1u << K;
    const uint32_t lowmask = (1u << lowb) - 1u, tmask = (1u << k) - 1u;
    uint32_t mb = 0;
    if constexpr (K <= 13) {
        if (lowb >= 1) {
            constexpr int IT = NE / (2 * NT);
            float4 v[IT];
#pragma unroll
            for (int i = 0; i < IT; ++i) {
                const uint32_t e = 2u * (i * NT + tid);
                const uint32_t j = e & tmask, u = e >> k;
                const uint64_t g = tb[u] | sbo[j >> lowb] | (j & lowmask);
                v[i] = ld_stream(reinterpret_cast<const float *>(st) + 2 * g);
            }
#pragma unroll
            for (int i = 0; i < IT; ++i)
                mb = max(mb, max(max(absbits(v[i].x), absbits(v[i].y)), max(absbits(v[i].z), absbits(v[i].w))));
            const int sh = tile_shift<NT>(mb, red, tid, target);
#pragma unroll
            for (int i = 0; i < IT; ++i) {
                const uint32_t pe = lpad(2u * (i * NT + tid));
                tile[pe] = C2<float>{ldexpf(v[i].x, sh), ldexpf(v[i].y, sh)};
                tile[pe + 1] = C2<float>{ldexpf(v[i].z, sh), ldexpf(v[i].w, sh)};
            }
            return sh;
        }
    }
    move_tile<float, K, true, NT>(st, tile, sbo, tb, k, lowb, tid);
    __syncthreads();
    for (uint32_t e = tid; e < NE; e += NT) {
        const C2<float> x = tile[lpad(e)];
        mb = max(mb, max(absbits(x.x), absbits(x.y)));
    }
    const int sh = tile_shift<NT>(mb, red, tid, target);
    if (sh != 0)
        for (uint32_t e = tid; e < NE; e += NT) {
            const C2<float> x = tile[lpad(e)];
            tile[lpad(e)] = C2<float>{ldexpf(x.x, sh), ldexpf(x.y, sh)};
        }
    return sh;
}

// ------------------------------------------------------------------------------------------ dense op
// Dense 2^F x 2^F complex op on workgroup-tile bits dq[0..F-1] (fused-form records and pass fusion),
// applied to the LDS tile on the matrix cores.  Real form Y = W X with W = [[Ur, -Ui], [Ui, Ur]]
// (2^(F+1) square) and the columns of X the 2^(K-F) groups of 2^F amplitudes (rows: Re x_0..,
// Im x_0..).  A wavefront handles 32 columns at a time (lane l owns column l % 32); the accumulator
// layout of a 32x32 output block is row i = 8 (v / 4) + 4 (l / 32) + v % 4 for accumulator v.
//
typedef _Float16 half8 __attribute__((ext_vector_type(8)));

// Split-FP16 matrix cores: W (scaled on the host by 2^wexp -- 0 for entries within [2^-4, 2], else
// so its largest entry is below 2^14 -- and split W 2^wexp = Wh + Wl) times a column block x scaled
// by a power of two s (per tile or per column, max |x s| < 2^15) and split x s = xh + xl in FP16;
// y = (Wh xh + Wh xl + Wl xh) / (2^wexp s) with FP32 accumulation.  The dropped product and the
// two-piece representations leave a relative error of about 2^-21 of the column's largest entry
// per product.  v_mfma_f32_32x32x16_f16: lane l holds row (A) / column (B) l % 32 and K entries
// 8 (l / 32) + 0..7 of each 16-deep K step.
//
// The real form is interleaved: real index 2c + p is component c, part p (0 re, 1 im), for rows and
// columns alike, so W[2a + p][2c + q] = (Ur, -Ui; Ui, Ur)[p][q] (a, c).  K step ks then holds
// components 8 ks .. 8 ks + 7, lane half h the four components 8 ks + 4 h + 0..3 as (re, im) pairs:
// one complex amplitude converts to one packed FP16 pair.  Accumulators v, v + 1 (v even) of 32-row
// tile rt are re, im of component 16 rt + 4 (v / 4) + 2 h + (v % 4) / 2 and store as one 8-byte pair.
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

typedef float float2v __attribute__((ext_vector_type(2)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));

// (re, im) = hi + lo in FP16 pairs (re, im both below 2^15 in magnitude): hi keeps the top 11
// significant bits (a mask, then an exact conversion; the FP16 subnormal grid below 2^-14), lo =
// (re, im) - hi -- exact in FP32 -- rounded to the nearest FP16, so the split does not shrink
// amplitudes on average and the norm does not drift over many dense ops.
__device__ __forceinline__ void split_pk(float re, float im, uint32_t &hi, uint32_t &lo) {
    const float rh = __uint_as_float(__float_as_uint(re) & 0xffffe000u);
    const float ih = __uint_as_float(__float_as_uint(im) & 0xffffe000u);
    hi = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(rh, ih));
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(float2v{re - rh, im - ih}, half2v));
}

template <int K, int F, int NWAVE, bool SWZ, bool PRE>
__device__ __forceinline__ void dense_apply_h(unsigned char *buf, const GroupDev &G, const float *__restrict__ wd,
                                              uint32_t wt) {
    constexpr int NRT = F == 5 ? 2 : 1;            // 32-row output tiles
    constexpr int D = 1 << F;
    constexpr int NKS = D / 8;                     // 16-deep K steps of 8 interleaved components
    constexpr uint32_t NB = (1u << (K - F)) / 32;
    const uint32_t lane = wt & 63u, wave = wt >> 6, h = lane >> 5, j = lane & 31u;
    const half8 *w8 = reinterpret_cast<const half8 *>(wd + G.w_off);
    half8 ah[NRT][NKS], al[NRT][NKS];
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) {
            ah[rt][ks] = w8[((rt * NKS + ks) * 2 + 0) * 64 + lane];
            al[rt][ks] = w8[((rt * NKS + ks) * 2 + 1) * 64 + lane];
        }
    ColOff<F, SWZ> coff;
    coff.init(G);
    // column b 32 + j sits at tile index dep(b 32) | dep(j) (dep: spread the bits over the positions
    // outside the unit's qubits; disjoint source bits spread to disjoint positions, and both tile
    // layouts are additive / XOR-linear over disjoint bits): the lane part once, the wave-uniform
    // part in scalar registers per step
    uint32_t dj = j;
#pragma unroll
    for (int i = 0; i < F; ++i) dj = ins0_32(dj, G.dqs[i]);
    const uint32_t pj = TileAddr<SWZ>::of(dj);
    // NP independent 32-column batches per step (b and b + NWAVE): one batch's MFMA chain overlaps
    // the other's LDS reads, split and stores within the wavefront (4-qubit units: the registers fit)
    constexpr int NP = (F == 4 && NB % (2u * NWAVE) == 0) ? 2 : 1;
    for (uint32_t b = __builtin_amdgcn_readfirstlane(wave); b < NB; b += NP * NWAVE) {
        // the lane half's component bit goes into the lane base, the rest of each component's
        // offset is uniform (scalar registers)
        uint32_t pbl[NP], pbs[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            uint32_t db = (b + p * NWAVE) * 32u;
#pragma unroll
            for (int i = 0; i < F; ++i) db = ins0_32(db, G.dqs[i]);
            const uint32_t pb = TileAddr<SWZ>::comb(pj, TileAddr<SWZ>::of(db));
            pbl[p] = TileAddr<SWZ>::comb(pb, h ? coff.lq[2] : 0u);
            pbs[p] = TileAddr<SWZ>::comb(pb, h ? coff.lq[1] : 0u);
        }
        C2<float> x[NP][NKS][4];
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
                for (int i = 0; i < 4; ++i) x[p][ks][i] = ldc(buf, TileAddr<SWZ>::comb(pbl[p], coff(8u * ks + i)));
        int sh[NP];   // pre-scaled tile: every component is already below 2^15
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            sh[p] = 0;
            if constexpr (!PRE) {
                uint32_t mb = 0;
#pragma unroll
                for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
                    for (int i = 0; i < 4; ++i) mb = max(mb, max(absbits(x[p][ks][i].x), absbits(x[p][ks][i].y)));
                mb = max(mb, (uint32_t)__shfl_xor((int)mb, 32));
                const int efield = (int)(mb >> 23);
                sh[p] = mb == 0 ? 0 : 15 - (efield == 0 ? -125 : efield - 126);   // column maximum < 2^15
            }
        }
        half8 bh[NP][NKS], bl[NP][NKS];
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int ks = 0; ks < NKS; ++ks) {
                u4v hv, lv;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float re = PRE ? x[p][ks][i].x : ldexpf(x[p][ks][i].x, sh[p]);
                    const float im = PRE ? x[p][ks][i].y : ldexpf(x[p][ks][i].y, sh[p]);
                    uint32_t hh, ll;
                    split_pk(re, im, hh, ll);
                    hv[i] = hh;
                    lv[i] = ll;
                }
                bh[p][ks] = __builtin_bit_cast(half8, hv);
                bl[p][ks] = __builtin_bit_cast(half8, lv);
            }
        f16v acc[NP][NRT];
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt) {
#pragma unroll
            for (int p = 0; p < NP; ++p) acc[p][rt] = f16v{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    acc[p][rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[rt][ks], bh[p][ks], acc[p][rt], 0, 0, 0);
                    acc[p][rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[rt][ks], bl[p][ks], acc[p][rt], 0, 0, 0);
                }
#pragma unroll
            for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
                for (int p = 0; p < NP; ++p)
                    acc[p][rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[rt][ks], bh[p][ks], acc[p][rt], 0, 0, 0);
        }
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int back = -(sh[p] + G.wexp);
            if (back != 0) {
#pragma unroll
                for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
                    for (int v = 0; v < 16; ++v) acc[p][rt][v] = ldexpf(acc[p][rt][v], back);
            }
#pragma unroll
            for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
                for (int v = 0; v < 16; v += 2) {
                    const uint32_t c = 16u * rt + 4u * (v / 4) + (v % 4) / 2;   // + 2 h: in pbs
                    stc(buf, TileAddr<SWZ>::comb(pbs[p], coff(c)), C2<float>{acc[p][rt][v], acc[p][rt][v + 1]});
                }
        }
    }
}

// One 32-column batch of a 4-qubit split-FP16 unit in registers: x[ks][i] is component
// 8 ks + 4 h + i of the lane's column; returns Y = W X (accumulator v, v + 1: re, im of component
// 4 (v / 4) + 2 h + (v % 4) / 2), unscaled.
template <bool PRE>
__device__ __forceinline__ f16v unit4_h(const C2<float> (&x)[2][4], const half8 (&ah)[2], const half8 (&al)[2],
                                        int wexp) {
    int sh = 0;   // pre-scaled tile: every component is already below 2^15
    if constexpr (!PRE) {
        uint32_t mb = 0;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int i = 0; i < 4; ++i) mb = max(mb, max(absbits(x[ks][i].x), absbits(x[ks][i].y)));
        mb = max(mb, (uint32_t)__shfl_xor((int)mb, 32));
        const int efield = (int)(mb >> 23);
        sh = mb == 0 ? 0 : 15 - (efield == 0 ? -125 : efield - 126);   // column maximum < 2^15
    }
    half8 bh[2], bl[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        u4v hv, lv;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint32_t hh, ll;
            split_pk(PRE ? x[ks][i].x : ldexpf(x[ks][i].x, sh), PRE ? x[ks][i].y : ldexpf(x[ks][i].y, sh), hh, ll);
            hv[i] = hh;
            lv[i] = ll;
        }
        bh[ks] = __builtin_bit_cast(half8, hv);
        bl[ks] = __builtin_bit_cast(half8, lv);
    }
    f16v acc = f16v{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[ks], bh[ks], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[ks], bl[ks], acc, 0, 0, 0);
    }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[ks], bh[ks], acc, 0, 0, 0);
    const int back = -(sh + wexp);
    if (back != 0) {
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[v] = ldexpf(acc[v], back);
    }
    return acc;
}

// 4-qubit split-FP16 unit on the swizzled tile with 16-byte LDS accesses (same arithmetic as
// dense_apply_h).  Tile bit 0 inside the unit (the host makes it the matrix's index bit 0):
// components 2c, 2c + 1 of a column share a 16-byte slot, two loads per K step.  Tile bit 0 outside:
// a step covers 64 columns, the lane's columns 2j' and 2j' + 1 share every slot, and the two
// 32-column batches of the step are those two halves.  Returns false (nothing done) when the tile
// has too few column steps for the wavefronts; the caller then uses dense_apply_h.
template <int K, int NWAVE, bool PRE>
__device__ __forceinline__ bool dense4_h128(unsigned char *buf, const GroupDev &G, const float *__restrict__ wd,
                                            uint32_t wt) {
    constexpr uint32_t NCOL = 1u << (K - 4);
    const bool inner = G.dq[0] == 0;
    const uint32_t nstep = inner ? NCOL / 32 : NCOL / 64;
    constexpr int NP = (NCOL / 32) % (2 * NWAVE) == 0 ? 2 : 1;   // batches per step, tile bit 0 inside
    if (inner ? (nstep % (NP * NWAVE) != 0) : (nstep % NWAVE != 0)) return false;
    const uint32_t lane = wt & 63u, wave = wt >> 6, h = lane >> 5, j = lane & 31u;
    const half8 *w8 = reinterpret_cast<const half8 *>(wd + G.w_off);
    half8 ah[2], al[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        ah[ks] = w8[(ks * 2 + 0) * 64 + lane];
        al[ks] = w8[(ks * 2 + 1) * 64 + lane];
    }
    ColOff<4, true> coff;
    coff.init(G);
    auto dep = [&](uint32_t c) {   // spread c over the tile bits outside the unit
#pragma unroll
        for (int i = 0; i < 4; ++i) c = ins0_32(c, G.dqs[i]);
        return c;
    };
    if (inner) {
        const uint32_t pj = swz(dep(j));
        for (uint32_t b = __builtin_amdgcn_readfirstlane(wave); b < nstep; b += NP * NWAVE) {
            uint32_t pbl[NP], pbs[NP];
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                const uint32_t pb = pj ^ swz(dep((b + p * NWAVE) * 32u));
                pbl[p] = pb ^ (h ? coff.lq[2] : 0u);
                pbs[p] = pb ^ (h ? coff.lq[1] : 0u);
            }
            C2<float> x[NP][2][4];
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int ks = 0; ks < 2; ++ks)
#pragma unroll
                    for (int i = 0; i < 4; i += 2) {   // components 8 ks + 4 h + i, + 1: one slot
                        const float4 v = *reinterpret_cast<const float4 *>(buf + (pbl[p] ^ coff(8u * ks + i)));
                        x[p][ks][i] = C2<float>{v.x, v.y};
                        x[p][ks][i + 1] = C2<float>{v.z, v.w};
                    }
            f16v acc[NP];
#pragma unroll
            for (int p = 0; p < NP; ++p) acc[p] = unit4_h<PRE>(x[p], ah, al, G.wexp);
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int v = 0; v < 16; v += 4)   // components 4 (v / 4) + 2 h, + 1: one slot
                    *reinterpret_cast<float4 *>(buf + (pbs[p] ^ coff(4u * (v / 4)))) =
                        make_float4(acc[p][v], acc[p][v + 1], acc[p][v + 2], acc[p][v + 3]);
        }
    } else {
        const uint32_t pj = swz(dep(2u * j));
        for (uint32_t b = __builtin_amdgcn_readfirstlane(wave); b < nstep; b += NWAVE) {
            const uint32_t pb = pj ^ swz(dep(b * 64u));
            const uint32_t pbl = pb ^ (h ? coff.lq[2] : 0u), pbs = pb ^ (h ? coff.lq[1] : 0u);
            C2<float> x[2][2][4];
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
#pragma unroll
                for (int i = 0; i < 4; ++i) {   // component 8 ks + 4 h + i of columns 2j', 2j' + 1
                    const float4 v = *reinterpret_cast<const float4 *>(buf + (pbl ^ coff(8u * ks + i)));
                    x[0][ks][i] = C2<float>{v.x, v.y};
                    x[1][ks][i] = C2<float>{v.z, v.w};
                }
            // both columns' results of a component go out as one 16-byte write (8-byte writes of
            // one column would all fall in the same half of their slots)
            f16v acc[2];
#pragma unroll
            for (int p = 0; p < 2; ++p) acc[p] = unit4_h<PRE>(x[p], ah, al, G.wexp);
#pragma unroll
            for (int v = 0; v < 16; v += 2) {
                const uint32_t c = 4u * (v / 4) + (v % 4) / 2;   // + 2 h: in pbs
                *reinterpret_cast<float4 *>(buf + (pbs ^ coff(c))) =
                    make_float4(acc[0][v], acc[0][v + 1], acc[1][v], acc[1][v + 1]);
            }
        }
    }
    return true;
}

// One dense group (type 1: FP32 matrix cores, 3: split FP16 with per-column scaling, 4: split FP16
// on a pre-scaled tile) by NWAVE wavefronts
template <int K, int NWAVE, bool SWZ>
__device__ __forceinline__ void dense_group(unsigned char *buf, const GroupDev &G, const float *__restrict__ wd,
                                            uint32_t wt) {
    if (G.type == 5) {
        Gauss4W gw;
        gauss4_weights(wd + G.w_off, wt, gw);
        dense_gauss4<K, NWAVE, SWZ>(buf, G, gw, wt);
    } else if (G.type == 1) {
        if (G.f == 5) dense_apply<K, 5, NWAVE, SWZ>(buf, G, wd, wt);
        else dense_apply<K, 4, NWAVE, SWZ>(buf, G, wd, wt);
    } else if (G.type == 4 && !SWZ) {
        if (G.f == 5) dense_apply_h<K, 5, NWAVE, SWZ, true>(buf, G, wd, wt);
        else dense_apply_h<K, 4, NWAVE, SWZ, true>(buf, G, wd, wt);
    } else {
        if (G.f == 5) dense_apply_h<K, 5, NWAVE, SWZ, false>(buf, G, wd, wt);
        else if (!SWZ || !dense4_h128<K, NWAVE, false>(buf, G, wd, wt)) dense_apply_h<K, 4, NWAVE, SWZ, false>(buf, G, wd, wt);
    }
}

// 4-qubit dense group only (the dense-only instance's register budget)
template <int K, int NWAVE, bool SWZ = false>
__device__ __forceinline__ void dense_group4(unsigned char *buf, const GroupDev &G, const float *__restrict__ wd,
                                             uint32_t wt) {
    if (G.type == 5) {
        Gauss4W gw;
        gauss4_weights(wd + G.w_off, wt, gw);
        dense_gauss4<K, NWAVE, SWZ>(buf, G, gw, wt);
    } else if (G.type == 1) {
        dense_apply<K, 4, NWAVE, SWZ>(buf, G, wd, wt);
    } else if (G.type == 4) {
        if (!SWZ || !dense4_h128<K, NWAVE, true>(buf, G, wd, wt)) dense_apply_h<K, 4, NWAVE, SWZ, true>(buf, G, wd, wt);
    } else {
        if (!SWZ || !dense4_h128<K, NWAVE, false>(buf, G, wd, wt)) dense_apply_h<K, 4, NWAVE, SWZ, false>(buf, G, wd, wt);
    }
}

// Dense-only passes keep the tile XOR-swizzled in LDS (swz: amplitudes 2i, 2i + 1 share one 16-byte
// slot, slot i ^ ((2i >> 5) & 15)), moved with 16-byte LDS accesses; every QSV2 tile has at least
// one contiguous low bit (k - m >= 1), so each thread's 16-byte HBM access is one slot.
// Load: returns the tile scaling shift (target > 0: largest component below 2^target).
template <int K, int NT>
__device__ __forceinline__ int tile_in_swz(const C2<float> *__restrict__ st, unsigned char *buf, const uint64_t *sbo,
                                           const uint64_t *tb, int k, int lowb, uint32_t tid, uint32_t *red,
                                           int target) {
    constexpr int IT = (1 << K) / (2 * NT);
    const uint32_t lowmask = (1u << lowb) - 1u, tmask = (1u << k) - 1u;
    float4 v[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const uint32_t e = 2u * (i * NT + tid);
        const uint32_t j = e & tmask, u = e >> k;
        v[i] = ld_stream(reinterpret_cast<const float *>(st) + 2 * (tb[u] | sbo[j >> lowb] | (j & lowmask)));
    }
    int sh = 0;
    if (target > 0) {
        uint32_t mb = 0;
#pragma unroll
        for (int i = 0; i < IT; ++i)
            mb = max(mb, max(max(absbits(v[i].x), absbits(v[i].y)), max(absbits(v[i].z), absbits(v[i].w))));
        sh = tile_shift<NT>(mb, red, tid, target);
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        float4 w = v[i];
        if (sh) w = make_float4(ldexpf(w.x, sh), ldexpf(w.y, sh), ldexpf(w.z, sh), ldexpf(w.w, sh));
        *reinterpret_cast<float4 *>(buf + swz(2u * (i * NT + tid))) = w;
    }
    return sh;
}
template <int K, int NT>
__device__ __forceinline__ void tile_out_swz(C2<float> *__restrict__ st, const unsigned char *buf, const uint64_t *sbo,
                                             const uint64_t *tb, int k, int lowb, uint32_t tid, int unscale) {
    constexpr int IT = (1 << K) / (2 * NT);
    constexpr int BATCH = IT < 4 ? IT : 4;   // registers: four 16-byte values in flight per step
    const uint32_t lowmask = (1u << lowb) - 1u, tmask = (1u << k) - 1u;
#pragma unroll
    for (int i0 = 0; i0 < IT; i0 += BATCH) {
        float4 w[BATCH];
#pragma unroll
        for (int i = 0; i < BATCH; ++i) {
            w[i] = *reinterpret_cast<const float4 *>(buf + swz(2u * ((i0 + i) * NT + tid)));
            if (unscale)
                w[i] = make_float4(ldexpf(w[i].x, unscale), ldexpf(w[i].y, unscale), ldexpf(w[i].z, unscale),
                                   ldexpf(w[i].w, unscale));
        }
#pragma unroll
        for (int i = 0; i < BATCH; ++i) {
            const uint32_t e = 2u * ((i0 + i) * NT + tid);
            const uint32_t j = e & tmask, u = e >> k;
            st_stream(reinterpret_cast<float *>(st) + 2 * (tb[u] | sbo[j >> lowb] | (j & lowmask)), w[i]);
        }
    }
}

// Diagonal table on a MUX2 register group: the host stores the phase of register r of slot s at
// entry r NSLOT + s, so for each register the threads of a wavefront read consecutive entries.
template <int R, uint32_t NSLOT>
__device__ __forceinline__ void table_regs(C2<float> (&a)[1 << R], const float *__restrict__ tab, uint32_t s) {
    const float2 *d = reinterpret_cast<const float2 *>(tab) + s;
#pragma unroll
    for (int r0 = 0; r0 < (1 << R); r0 += 8) {   // eight table reads in flight, few registers
        float2 w[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) w[r] = d[(r0 + r) * NSLOT];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const C2<float> v = a[r0 + r];
            a[r0 + r] = C2<float>{fmaf(v.x, w[r].x, -v.y * w[r].y), fmaf(v.x, w[r].y, v.y * w[r].x)};
        }
    }
}

// One workgroup tile of one pass: `tile` selects the 2^(K-k) QSV2 tiles the workgroup owns.
// DN: the pass has dense groups or diagonal tables; D4: dense-only instance (every group a dense
// 4-qubit unit, NT threads, no MUX2 register groups compiled in).
template <typename T, int K, bool DN, bool D4 = false, int NT = kNT>
__device__ __forceinline__ void pass_tile(C2<T> *__restrict__ st, const PassDev *__restrict__ pd,
                                          const GroupDev *__restrict__ gd, const OpDev<T> *__restrict__ od,
                                          const float *__restrict__ wd, uint64_t tile_id, unsigned char *smem) {
    constexpr uint32_t NE = 1u << K;
    constexpr uint32_t NEP = NE + (NE >> 5);
    C2<T> *tile = reinterpret_cast<C2<T> *>(smem);
    uint64_t *sbo = reinterpret_cast<uint64_t *>(smem + NEP * sizeof(C2<T>));
    const int k = pd->k, m = pd->m;
    const int sub = K - k;
    uint64_t *tb = sbo + (1u << m);
    const uint32_t tid = threadIdx.x;

    for (uint32_t s = tid; s < (1u << m); s += NT) {
        uint64_t o = 0;
        for (int i = 0; i < m; ++i)
            if ((s >> i) & 1u) o |= uint64_t(1) << pd->h[i];
        sbo[s] = o;
    }
    for (uint32_t u = tid; u < (1u << sub); u += NT) {
        uint64_t x = ((tile_id << sub) | u) << (k - m);
        for (int i = 0; i < m; ++i) x = ins0_64(x, pd->hs[i]);
        tb[u] = x;
    }
    __syncthreads();
    const int lowb = k - m;
    int tsh = 0;   // tile scaling shift (passes with pre-scaled dense groups)
    if constexpr (D4 && sizeof(T) == 4) {
        tsh = tile_in_swz<K, NT>(reinterpret_cast<const C2<float> *>(st), smem, sbo, tb, k, lowb, tid,
                                 reinterpret_cast<uint32_t *>(tb + (1u << sub)), pd->tscale);
    } else if constexpr (DN && sizeof(T) == 4) {
        if (pd->tscale > 0)
            tsh = load_tile_scaled<K, NT>(reinterpret_cast<C2<float> *>(st), reinterpret_cast<C2<float> *>(tile), sbo,
                                          tb, k, lowb, tid, reinterpret_cast<uint32_t *>(tb + (1u << sub)), pd->tscale);
        else
            move_tile<T, K, true, NT>(st, tile, sbo, tb, k, lowb, tid);
    } else {
        move_tile<T, K, true, NT>(st, tile, sbo, tb, k, lowb, tid);
    }
    __syncthreads();

    constexpr int RR = kR;   // register bits of a MUX2 group
    constexpr int NR = 1 << RR;
    constexpr uint32_t NSLOT = NE >> RR;
    const int ng = pd->ngroups, g0 = pd->group0;
    for (int g = 0; g < ng; ++g) {
        const GroupDev &G = gd[g0 + g];
        if constexpr (D4 && sizeof(T) == 4) {
            // dense-only passes: no MUX2 path compiled in
            dense_group4<K, NT / 64, true>(smem, G, wd, tid);
            __syncthreads();
            continue;
        } else if constexpr (DN && sizeof(T) == 4) {
            if (G.type == 6) {
                diag_table<K, NT>(reinterpret_cast<C2<float> *>(smem), wd + G.w_off, k, tid);
                __syncthreads();
                continue;
            }
            if (G.type != 0) {
                dense_group<K, NT / 64, false>(smem, G, wd, tid);
                __syncthreads();
                continue;
            }
        }
        uint32_t poff[NR];   // padded LDS offset of register r (disjoint bits: pad(a|b) = pad(a)+pad(b))
        poff[0] = 0;
#pragma unroll
        for (int q = 0; q < RR; ++q) {
            const uint32_t bq = lpad(1u << G.b[q]);
#pragma unroll
            for (int r = (1 << q); r < (2 << q); ++r) poff[r] = poff[r - (1 << q)] + bq;
        }
        const int nops = G.nops, o0 = G.op0;
#pragma unroll 1
        for (uint32_t s = tid; s < NSLOT; s += NT) {
            uint32_t base = s;
#pragma unroll
            for (int q = 0; q < RR; ++q) base = ins0_32(base, G.bs[q]);
            const uint32_t pb = lpad(base);
            C2<T> a[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) a[r] = tile[pb + poff[r]];
#pragma unroll 1
            for (int o = 0; o < nops; ++o) {
                const OpDev<T> &op = od[o0 + o];
                if constexpr (DN && sizeof(T) == 4) {
                    if (op.kind == K_TABLE) {
                        table_regs<RR, NSLOT>(reinterpret_cast<C2<float>(&)[NR]>(a), wd + op.aux, s);
                        continue;
                    }
                }
                const bool ts = (base & op.cmt) == op.cmt;
                apply_op<T, RR>(a, op, ts);
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) tile[pb + poff[r]] = a[r];
        }
        __syncthreads();
    }
    if constexpr (D4 && sizeof(T) == 4) {
        // the store addresses are recomputed here rather than kept live (as common subexpressions
        // of the load addresses) through the dense groups, which would spill registers
        uint32_t tt = tid;
        asm volatile("" : "+v"(tt));
        tile_out_swz<K, NT>(reinterpret_cast<C2<float> *>(st), smem, sbo, tb, k, lowb, tt, -tsh);
    }
    else
        move_tile<T, K, false, NT>(st, tile, sbo, tb, k, lowb, tid, -tsh);
}

template <typename T, int K, bool DN>
__global__ __launch_bounds__(kNT, (DN && K < 14) ? 2 : 1) void k_pass(C2<T> *__restrict__ st, const PassDev *__restrict__ pd,
                                              const GroupDev *__restrict__ gd,
                                              const OpDev<T> *__restrict__ od, const float *__restrict__ wd) {
    extern __shared__ __align__(16) unsigned char smem[];
    pass_tile<T, K, DN>(st, pd, gd, od, wd, tile_order(blockIdx.x, pd), smem);
}

// Dense-only passes with 4-qubit units: 512 threads and at most 128 VGPRs (90 used), so two workgroups
// give four waves per SIMD to hide one wave's LDS -> split -> MFMA -> LDS chain behind the others.
template <int K>
__global__ __launch_bounds__(kNTD4, 4) void k_pass_d4(C2<float> *__restrict__ st, const PassDev *__restrict__ pd,
                                                      const GroupDev *__restrict__ gd,
                                                      const OpDev<float> *__restrict__ od, const float *__restrict__ wd) {
    extern __shared__ __align__(16) unsigned char smem[];
    pass_tile<float, K, true, true, kNTD4>(st, pd, gd, od, wd, tile_order(blockIdx.x, pd), smem);
}

// Continuous batching: one launch runs one pass of each of several states (the items, one kernel
// shape per launch); workgroup b runs tile b - items[i].tile0 of the item i whose tile range holds
// b, with the same device code as k_pass, so every state gets the same arithmetic alone and in
// the batch.
template <int K, bool DN>
__global__ __launch_bounds__(kNT, (DN && K < 14) ? 2 : 1) void k_pass_batch(const BatchItem *__restrict__ items, int n_items) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint64_t b = blockIdx.x;
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {   // last item with tile0 <= b
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].tile0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const BatchItem &it = items[__builtin_amdgcn_readfirstlane(lo)];   // uniform: scalar loads
    pass_tile<float, K, DN>(reinterpret_cast<C2<float> *>(it.st), it.pd, it.gd,
                            reinterpret_cast<const OpDev<float> *>(it.od), it.wd, tile_order(b - it.tile0, it.pd), smem);
}


// One workgroup runs a whole small state (<= 2^kSmallMaxBits amplitudes), op by op, directly
// from the MUX2 definition.  Also the independent cross-check of k_pass in the tests.
template <typename T>
__global__ __launch_bounds__(kNT) void k_pass_small(C2<T> *__restrict__ st, const PassDev *__restrict__ pd,
                                                    const OpDev<T> *__restrict__ od,
                                                    const GroupDev *__restrict__ gd, const float *__restrict__ wd) {
    __shared__ C2<T> tile[1u << kSmallMaxBits];
    const int k = pd->k, m = pd->m, nb = pd->nbits;
    const int lowb = k - m;
    const uint32_t NE = 1u << nb, tmask = (1u << k) - 1u, lowmask = (1u << lowb) - 1u;
    const uint32_t tid = threadIdx.x;
    auto gidx = [&](uint32_t e) -> uint64_t {
        const uint32_t j = e & tmask, u = e >> k;
        uint64_t x = uint64_t(u) << lowb;
        for (int i = 0; i < m; ++i) x = ins0_64(x, pd->hs[i]);
        uint64_t o = j & lowmask;
        for (int i = 0; i < m; ++i)
            if ((j >> (lowb + i)) & 1u) o |= uint64_t(1) << pd->h[i];
        return x | o;
    };
    for (uint32_t e = tid; e < NE; e += kNT) tile[e] = st[gidx(e)];
    __syncthreads();
    const int nops = pd->nops, o0 = pd->op0;
    for (int o = 0; o < nops; ++o) {
        const OpDev<T> &op = od[o0 + o];
        if constexpr (sizeof(T) == 4) {
            if (op.kind == K_DENSE) {
                // y[row] = sum_c U[row][c] x[c] over the op's qubits; NE <= 2 * kNT
                const GroupDev &G = gd[op.cmask];
                const int f = G.f;
                const uint32_t D = 1u << f;
                const float *U = wd + G.w_off;
                uint32_t qm = 0;
                for (int i = 0; i < f; ++i) qm |= 1u << G.dq[i];
                C2<T> y[2];
                for (int r = 0; r < 2; ++r) {
                    const uint32_t e = tid + r * kNT;
                    if (e >= NE) break;
                    uint32_t row = 0;
                    for (int i = 0; i < f; ++i) row |= ((e >> G.dq[i]) & 1u) << i;
                    const uint32_t base = e & ~qm;
                    float ar = 0.f, ai = 0.f;
                    for (uint32_t c = 0; c < D; ++c) {
                        uint32_t idx = base;
                        for (int i = 0; i < f; ++i)
                            if ((c >> i) & 1u) idx |= 1u << G.dq[i];
                        const float ur = U[2 * (row * D + c)], ui = U[2 * (row * D + c) + 1];
                        const C2<T> x = tile[idx];
                        ar = fmaf(ur, x.x, fmaf(-ui, x.y, ar));
                        ai = fmaf(ur, x.y, fmaf(ui, x.x, ai));
                    }
                    y[r] = C2<T>{ar, ai};
                }
                __syncthreads();
                for (int r = 0; r < 2; ++r) {
                    const uint32_t e = tid + r * kNT;
                    if (e < NE) tile[e] = y[r];
                }
                __syncthreads();
                continue;
            }
        }
        const int t = op.t;
        const uint32_t cm = uint32_t(op.cmask);
        for (uint32_t p = tid; p < NE / 2; p += kNT) {
            const uint32_t i0 = ins0_32(p, t), i1 = i0 | (1u << t);
            const bool sel = cm != 0 && ((i0 & tmask & cm) == cm);
            T mm[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) mm[e] = sel ? op.m1[e] : op.m0[e];
            C2<T> x0 = tile[i0], x1 = tile[i1];
            mat2(mm, x0, x1);
            tile[i0] = x0;
            tile[i1] = x1;
        }
        __syncthreads();
    }
    for (uint32_t e = tid; e < NE; e += kNT) st[gidx(e)] = tile[e];
}

size_t pass_lds_bytes(int K, int m, int k, int elem_bytes, bool small, int nbits) {
    if (small) return 0;
    const size_t nep = (size_t(1) << K) + ((size_t(1) << K) >> 5);
    return nep * elem_bytes + ((size_t(1) << m) + (size_t(1) << (K - k))) * sizeof(uint64_t) + 64;
}

template <typename T>
static hipError_t launch_pass_t(int K, bool small, void *st, const PassDev *pd, const GroupDev *gd,
                                const OpDev<T> *od, const float *wd, uint32_t grid, size_t lds, hipStream_t s) {
    C2<T> *p = reinterpret_cast<C2<T> *>(st);
    if (small) {
        hipLaunchKernelGGL(k_pass_small<T>, dim3(1), dim3(kNT), 0, s, p, pd, od, gd, wd);
        return hipGetLastError();
    }
    if constexpr (sizeof(T) == 4) {
        if (wd) {   // pass with dense (fused-form) groups: the MFMA instance
            switch (K) {
            case 10: hipLaunchKernelGGL((k_pass<T, 10, true>), dim3(grid), dim3(kNT), lds, s, p, pd, gd, od, wd); break;
            case 11: hipLaunchKernelGGL((k_pass<T, 11, true>), dim3(grid), dim3(kNT), lds, s, p, pd, gd, od, wd); break;
            case 12: hipLaunchKernelGGL((k_pass<T, 12, true>), dim3(grid), dim3(kNT), lds, s, p, pd, gd, od, wd); break;
            case 13: hipLaunchKernelGGL((k_pass<T, 13, true>), dim3(grid), dim3(kNT), lds, s, p, pd, gd, od, wd); break;
            case 14: hipLaunchKernelGGL((k_pass<T, 14, true>), dim3(grid), dim3(kNT), lds, s, p, pd, gd, od, wd); break;
            default: return hipErrorInvalidValue;
            }
            return hipGetLastError();
        }
    }
    switch (K) {
    case 10: hipLaunchKernelGGL((k_pass<T, 10, false>), dim3(grid), dim3(kNT), lds, s, p, pd, gd, od, wd); break;
    case 11: hipLaunchKernelGGL((k_pass<T, 11, false>), dim3(grid), dim3(kNT), lds, s, p, pd, gd, od, wd); break;
    case 12: hipLaunchKernelGGL((k_pass<T, 12, false>), dim3(grid), dim3(kNT), lds, s, p, pd, gd, od, wd); break;
    case 13: hipLaunchKernelGGL((k_pass<T, 13, false>), dim3(grid), dim3(kNT), lds, s, p, pd, gd, od, wd); break;
    case 14:
        if constexpr (sizeof(T) == 4) {
            hipLaunchKernelGGL((k_pass<T, 14, false>), dim3(grid), dim3(kNT), lds, s, p, pd, gd, od, wd);
            break;
        } else {
            return hipErrorInvalidValue;
        }
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_pass_c64(int K, bool small, void *st, const PassDev *pd, const GroupDev *gd,
                           const OpDev<float> *od, const float *wd, uint32_t grid, size_t lds, hipStream_t s) {
    return launch_pass_t<float>(K, small, st, pd, gd, od, wd, grid, lds, s);
}
hipError_t launch_pass_d4(int K, void *st, const PassDev *pd, const GroupDev *gd, const OpDev<float> *od,
                          const float *wd, uint32_t grid, size_t lds, hipStream_t s) {
    C2<float> *p = reinterpret_cast<C2<float> *>(st);
    switch (K) {
    case 12: hipLaunchKernelGGL((k_pass_d4<12>), dim3(grid), dim3(kNTD4), lds, s, p, pd, gd, od, wd); break;
    case 13: hipLaunchKernelGGL((k_pass_d4<13>), dim3(grid), dim3(kNTD4), lds, s, p, pd, gd, od, wd); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
// batched launches of dense-only passes of 4-qubit units (the 512-thread kernel's passes)
template <int K>
__global__ __launch_bounds__(kNTD4, 4) void k_pass_batch_d4(const BatchItem *__restrict__ items, int n_items) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint64_t b = blockIdx.x;
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].tile0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const BatchItem &it = items[__builtin_amdgcn_readfirstlane(lo)];   // uniform: scalar loads
    pass_tile<float, K, true, true, kNTD4>(reinterpret_cast<C2<float> *>(it.st), it.pd, it.gd,
                                           reinterpret_cast<const OpDev<float> *>(it.od), it.wd,
                                           tile_order(b - it.tile0, it.pd), smem);
}

hipError_t launch_pass_batch(int K, bool dense, bool d4, const BatchItem *items, int n_items, uint32_t grid,
                             size_t lds, hipStream_t s) {
    if (d4) {
        switch (K) {
        case 12: hipLaunchKernelGGL((k_pass_batch_d4<12>), dim3(grid), dim3(kNTD4), lds, s, items, n_items); break;
        case 13: hipLaunchKernelGGL((k_pass_batch_d4<13>), dim3(grid), dim3(kNTD4), lds, s, items, n_items); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (K * 2 + (dense ? 1 : 0)) {
    case 20: hipLaunchKernelGGL((k_pass_batch<10, false>), dim3(grid), dim3(kNT), lds, s, items, n_items); break;
    case 21: hipLaunchKernelGGL((k_pass_batch<10, true>), dim3(grid), dim3(kNT), lds, s, items, n_items); break;
    case 22: hipLaunchKernelGGL((k_pass_batch<11, false>), dim3(grid), dim3(kNT), lds, s, items, n_items); break;
    case 23: hipLaunchKernelGGL((k_pass_batch<11, true>), dim3(grid), dim3(kNT), lds, s, items, n_items); break;
    case 24: hipLaunchKernelGGL((k_pass_batch<12, false>), dim3(grid), dim3(kNT), lds, s, items, n_items); break;
    case 25: hipLaunchKernelGGL((k_pass_batch<12, true>), dim3(grid), dim3(kNT), lds, s, items, n_items); break;
    case 26: hipLaunchKernelGGL((k_pass_batch<13, false>), dim3(grid), dim3(kNT), lds, s, items, n_items); break;
    case 27: hipLaunchKernelGGL((k_pass_batch<13, true>), dim3(grid), dim3(kNT), lds, s, items, n_items); break;
    case 28: hipLaunchKernelGGL((k_pass_batch<14, false>), dim3(grid), dim3(kNT), lds, s, items, n_items); break;
    case 29: hipLaunchKernelGGL((k_pass_batch<14, true>), dim3(grid), dim3(kNT), lds, s, items, n_items); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_pass_c128(int K, bool small, void *st, const PassDev *pd, const GroupDev *gd,
                            const OpDev<double> *od, uint32_t grid, size_t lds, hipStream_t s) {
    return launch_pass_t<double>(K, small, st, pd, gd, od, nullptr, grid, lds, s);
}

// ------------------------------------------------------------------------------------------ permute
// Workgroup tile: the new-index bits V (v of them: new bits 0..c-1, the new positions of old bits
// 0..c-1, then more new bits up to v); the other n - v bits are fixed per workgroup.  Reads walk
// the tile in old-index order (2^c contiguous old amplitudes per run), writes in new-index order.
template <typename T>
__global__ __launch_bounds__(256) void k_permute(C2<T> *__restrict__ dst, const C2<T> *__restrict__ src,
                                                 const PermDev *__restrict__ pd) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int c = pd->c, v = pd->v, nfix = pd->nfix;
    const uint32_t NE = 1u << v, cm = (1u << c) - 1u;
    // the descriptor's per-element tables go to LDS first (after the tile), so the address of each
    // HBM access is two LDS reads away instead of a dependent global load
    C2<T> *tile = reinterpret_cast<C2<T> *>(smem);
    uint64_t *t_oldhi = reinterpret_cast<uint64_t *>(smem + (size_t(NE) + (NE >> 4)) * sizeof(C2<T>));
    uint64_t *t_newhi = t_oldhi + 64;
    uint16_t *t_rdlo = reinterpret_cast<uint16_t *>(t_newhi + 64);
    uint16_t *t_rdhi = t_rdlo + 64;
    if (threadIdx.x < 64) {
        t_oldhi[threadIdx.x] = pd->oldhi[threadIdx.x];
        t_newhi[threadIdx.x] = pd->newhi[threadIdx.x];
        t_rdlo[threadIdx.x] = pd->rdlo[threadIdx.x];
        t_rdhi[threadIdx.x] = pd->rdhi[threadIdx.x];
    }
    uint64_t nb = 0, ob = 0;
    const uint64_t wg = blockIdx.x;
    for (int i = 0; i < nfix; ++i) {
        if ((wg >> i) & 1u) {
            nb |= uint64_t(1) << pd->fixnew[i];
            ob |= uint64_t(1) << pd->fixold[i];
        }
    }
    __syncthreads();
    if (sizeof(T) == 4 && c >= 1) {
        // complex64: adjacent element pairs (16 bytes) on both sides; LDS pads 2 elements per 32 so
        // new-order pairs stay 16-byte aligned
        // every 16-byte load of the tile (at most 8 per thread: v <= 12) is issued before the first
        // LDS write, and both sides stream past the caches (each amplitude is moved once)
        auto lp = [](uint32_t e) { return e + ((e >> 5) << 1); };
        const uint32_t b0 = uint32_t(t_rdlo[1]);    // LDS position bit of old bit 0
        float4 r[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t q = 2 * threadIdx.x + 512u * i;
            if (q < NE) r[i] = ld_stream(reinterpret_cast<const float *>(src) + 2 * (ob | (q & cm) | t_oldhi[q >> c]));
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t q = 2 * threadIdx.x + 512u * i;
            if (q < NE) {
                const uint32_t pos = uint32_t(t_rdlo[q & 63u]) | uint32_t(t_rdhi[q >> 6]);
                tile[lp(pos)] = C2<T>{T(r[i].x), T(r[i].y)};
                tile[lp(pos | b0)] = C2<T>{T(r[i].z), T(r[i].w)};
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t l = 2 * threadIdx.x + 512u * i;
            if (l < NE) {
                const C2<T> a0 = tile[lp(l)], a1 = tile[lp(l) + 1];
                st_stream(reinterpret_cast<float *>(dst) + 2 * (nb | (l & cm) | t_newhi[l >> c]),
                          make_float4(float(a0.x), float(a0.y), float(a1.x), float(a1.y)));
            }
        }
        return;
    }
    for (uint32_t q = threadIdx.x; q < NE; q += 256) {
        const uint32_t pos = uint32_t(t_rdlo[q & 63u]) | uint32_t(t_rdhi[q >> 6]);
        tile[lpad(pos)] = src[ob | (q & cm) | t_oldhi[q >> c]];
    }
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < NE; l += 256) dst[nb | (l & cm) | t_newhi[l >> c]] = tile[lpad(l)];
}

hipError_t launch_permute(int dtype, void *dst, const void *src, const PermDev *pd, uint32_t grid,
                          size_t lds, hipStream_t s) {
    if (dtype == 0)
        hipLaunchKernelGGL(k_permute<float>, dim3(grid), dim3(256), lds, s, (C2<float> *)dst,
                           (const C2<float> *)src, pd);
    else
        hipLaunchKernelGGL(k_permute<double>, dim3(grid), dim3(256), lds, s, (C2<double> *)dst,
                           (const C2<double> *)src, pd);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------ misc
__global__ void k_swap_words(uint4 *__restrict__ a, uint4 *__restrict__ b, int64_t n16) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += int64_t(gridDim.x) * blockDim.x) {
        const uint4 x = a[i], y = b[i];
        a[i] = y;
        b[i] = x;
    }
}

static uint32_t grid_for(int64_t n, int per_thread) {
    int64_t g = (n + int64_t(256) * per_thread - 1) / (int64_t(256) * per_thread);
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return uint32_t(g);
}

hipError_t launch_swap_words(void *a, void *b, int64_t offa, int64_t offb, int64_t nwords, hipStream_t s) {
    uint4 *pa = reinterpret_cast<uint4 *>(a) + offa * 4, *pb = reinterpret_cast<uint4 *>(b) + offb * 4;
    const int64_t n16 = nwords * 4;
    if (n16 <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_swap_words, dim3(grid_for(n16, 4)), dim3(256), 0, s, pa, pb, n16);
    return hipGetLastError();
}

template <typename T>
__global__ void k_zero(C2<T> *__restrict__ st, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        st[i] = C2<T>{i == 0 ? T(1) : T(0), T(0)};
}

hipError_t launch_zero(int dtype, void *st, int64_t n, hipStream_t s) {
    if (dtype == 0)
        hipLaunchKernelGGL(k_zero<float>, dim3(grid_for(n, 8)), dim3(256), 0, s, (C2<float> *)st, n);
    else
        hipLaunchKernelGGL(k_zero<double>, dim3(grid_for(n, 8)), dim3(256), 0, s, (C2<double> *)st, n);
    return hipGetLastError();
}

template <typename T>
__global__ void k_product(C2<T> *__restrict__ st, int64_t n, const double *__restrict__ tlo,
                          const double *__restrict__ thi, int nlo) {
    const int64_t lmask = (int64_t(1) << nlo) - 1;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const double ar = tlo[2 * (i & lmask)], ai = tlo[2 * (i & lmask) + 1];
        const double br = thi[2 * (i >> nlo)], bi = thi[2 * (i >> nlo) + 1];
        st[i] = C2<T>{T(ar * br - ai * bi), T(ar * bi + ai * br)};
    }
}

hipError_t launch_product(int dtype, void *st, int32_t n, const double *tlo, const double *thi, int32_t nlo,
                          hipStream_t s) {
    const int64_t N = int64_t(1) << n;
    if (dtype == 0)
        hipLaunchKernelGGL(k_product<float>, dim3(grid_for(N, 8)), dim3(256), 0, s, (C2<float> *)st, N, tlo, thi, nlo);
    else
        hipLaunchKernelGGL(k_product<double>, dim3(grid_for(N, 8)), dim3(256), 0, s, (C2<double> *)st, N, tlo, thi, nlo);
    return hipGetLastError();
}

template <typename TD, typename TS>
__global__ void k_convert(C2<TD> *__restrict__ d, const C2<TS> *__restrict__ s, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        d[i] = C2<TD>{TD(s[i].x), TD(s[i].y)};
}

// Same-type copy at the HBM roof: one 16-byte non-temporal load and store per thread and launch
// round, one-shot workgroups up to 2^22 of them (measured on 8 GiB: one access per thread in flight
// beats 2..16 per thread and grid-stride loops); `tail` copies a last odd 8-byte element.
__global__ __launch_bounds__(256) void k_copy16(f4v *__restrict__ d, const f4v *__restrict__ s, int64_t n16,
                                                uint2 *__restrict__ dt, const uint2 *__restrict__ st, int tail) {
    const int64_t stride = int64_t(gridDim.x) * 256;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n16; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(s + i), d + i);
    if (tail && blockIdx.x == 0 && threadIdx.x == 0) *dt = *st;
}

hipError_t launch_convert(void *dst, int dt_dst, const void *src, int dt_src, int64_t n, hipStream_t s) {
    const uint32_t g = grid_for(n, 8);
    if (dt_dst == dt_src) {
        const int64_t bytes = n * (dt_src == 0 ? 8 : 16), n16 = bytes / 16;
        const int tail = int(bytes & 15) ? 1 : 0;
        int64_t gb = (n16 + 255) / 256;
        if (gb > (int64_t(1) << 22)) gb = int64_t(1) << 22;
        if (gb < 1) gb = 1;
        hipLaunchKernelGGL(k_copy16, dim3(uint32_t(gb)), dim3(256), 0, s, (f4v *)dst, (const f4v *)src, n16,
                           (uint2 *)((char *)dst + n16 * 16), (const uint2 *)((const char *)src + n16 * 16), tail);
    } else if (dt_dst == 0)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(256), 0, s, (C2<float> *)dst, (const C2<double> *)src, n);
    else
        hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(256), 0, s, (C2<double> *)dst, (const C2<float> *)src, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------ FP64 reductions
// Marginals: workgroup w walks chunks w, w + G, ... of 2^c amplitudes in order.  Inside a chunk the
// 256 threads own the low 8 index bits; per value of the marginal bits in [8, c) (beta) each thread
// sums its amplitudes over the other bits in [8, c) (gamma) in order, then a fixed-order tree over
// the thread bits that are not marginal bits folds the threads of each bin.  All orders are fixed,
// so the result is bit-reproducible for a given grid.
template <typename T>
__global__ __launch_bounds__(256) void k_marg(const C2<T> *__restrict__ st, const MargCfg *__restrict__ cfg,
                                              double *__restrict__ part) {
    __shared__ double red[256];
    __shared__ double acc[256];
    const int c = cfg->c, low8 = cfg->low8, nlow = cfg->nlow;
    const int nsa = cfg->nsa, nra = cfg->nra, nsb = cfg->nsb, nrb = cfg->nrb, nshi = cfg->nshi;
    const uint32_t t = threadIdx.x;
    const uint32_t nbin = 1u << nlow;
    for (uint32_t x = t; x < 256; x += 256) acc[x] = 0.0;
    const bool active = t < (1u << low8);
    uint32_t ra_lo = 0;   // thread has no ra bits set below the current tree level
    uint32_t xa = 0;
    for (int i = 0; i < nsa; ++i)
        if ((t >> cfg->sa[i]) & 1u) xa |= 1u << cfg->sa_x[i];
    const uint64_t nchunk = uint64_t(1) << (cfg->n - c);
    __syncthreads();
    for (uint64_t ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
        uint32_t xh = 0;
        for (int i = 0; i < nshi; ++i)
            if ((ch >> cfg->shi[i]) & 1u) xh |= 1u << cfg->shi_x[i];
        const uint64_t cbase = ch << c;
        for (uint32_t beta = 0; beta < (1u << nsb); ++beta) {
            uint32_t wb = 0, xb = 0;
            for (int i = 0; i < nsb; ++i)
                if ((beta >> i) & 1u) {
                    wb |= 1u << cfg->sb[i];
                    xb |= 1u << cfg->sb_x[i];
                }
            double s = 0.0;
            if (active) {
                for (uint32_t gam = 0; gam < (1u << nrb); ++gam) {
                    uint32_t w = t | wb;
                    for (int i = 0; i < nrb; ++i)
                        if ((gam >> i) & 1u) w |= 1u << cfg->rb[i];
                    const C2<T> a = st[cbase | w];
                    s = fma(double(a.x), double(a.x), fma(double(a.y), double(a.y), s));
                }
            }
            red[t] = s;
            __syncthreads();
            ra_lo = 0;
            for (int i = 0; i < nra; ++i) {
                const uint32_t bit = 1u << cfg->ra[i];
                if (active && !(t & bit) && !(t & ra_lo)) red[t] += red[t | bit];
                ra_lo |= bit;
                __syncthreads();
            }
            // threads with no ra bit set now hold their bin's sum
            if (active && !(t & ra_lo)) acc[xa | xb | xh] += red[t];
            __syncthreads();
        }
    }
    for (uint32_t x = t; x < nbin; x += 256) part[uint64_t(blockIdx.x) * nbin + x] = acc[x];
}

__global__ __launch_bounds__(256) void k_fold(const double *__restrict__ part, int64_t G, int width,
                                              double *__restrict__ out, int op_max_mask) {
    // out[x] = sum_g part[g * width + x] (or max over g where bit x of op_max_mask is set), g in order
    for (int x = threadIdx.x; x < width; x += 256) {
        const bool is_max = (op_max_mask >> x) & 1;
        double s = is_max ? 0.0 : 0.0;
        for (int64_t g = 0; g < G; ++g) {
            const double v = part[g * width + x];
            s = is_max ? fmax(s, v) : s + v;
        }
        out[x] = s;
    }
}

hipError_t launch_marginals(int dtype, const void *st, const MargCfg *cfg_dev, int nbin, double *part,
                            uint32_t grid, double *out, hipStream_t s) {
    if (dtype == 0)
        hipLaunchKernelGGL(k_marg<float>, dim3(grid), dim3(256), 0, s, (const C2<float> *)st, cfg_dev, part);
    else
        hipLaunchKernelGGL(k_marg<double>, dim3(grid), dim3(256), 0, s, (const C2<double> *)st, cfg_dev, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(256), 0, s, part, int64_t(grid), nbin, out, 0);
    return hipGetLastError();
}

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_fid1(const C2<TA> *__restrict__ a, const C2<TB> *__restrict__ b, int64_t n,
                                              double *__restrict__ part) {
    __shared__ double red[4][256];
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
        s0 = fma(ar, br, fma(ai, bi, s0));       // Re <a|b>
        s1 = fma(ar, bi, fma(-ai, br, s1));      // Im <a|b>
        s2 = fma(ar, ar, fma(ai, ai, s2));
        s3 = fma(br, br, fma(bi, bi, s3));
    }
    const int t = threadIdx.x;
    red[0][t] = s0; red[1][t] = s1; red[2][t] = s2; red[3][t] = s3;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int q = 0; q < 4; ++q) red[q][t] += red[q][t + w];
        __syncthreads();
    }
    if (t < 4) part[uint64_t(blockIdx.x) * 4 + t] = red[t][0];
}

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_fid2(const C2<TA> *__restrict__ a, const C2<TB> *__restrict__ b, int64_t n,
                                              const double *__restrict__ coef, double *__restrict__ part) {
    __shared__ double rs[256], rm[256];
    // coef: ph.re, ph.im, 1/|a|, 1/|b|;  e = b/|b| - ph a/|a|
    const double pr = coef[0], pi = coef[1], ia = coef[2], ib = coef[3];
    double d2 = 0, mx = 0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const double ar = a[i].x * ia, ai = a[i].y * ia;
        const double er = b[i].x * ib - (pr * ar - pi * ai);
        const double ei = b[i].y * ib - (pr * ai + pi * ar);
        const double m2 = er * er + ei * ei;
        d2 += m2;
        mx = fmax(mx, m2);
    }
    const int t = threadIdx.x;
    rs[t] = d2;
    rm[t] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            rs[t] += rs[t + w];
            rm[t] = fmax(rm[t], rm[t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        part[uint64_t(blockIdx.x) * 2] = rs[0];
        part[uint64_t(blockIdx.x) * 2 + 1] = rm[0];
    }
}

hipError_t launch_fid_pass1(int dta, const void *a, int dtb, const void *b, int64_t n, double *part, uint32_t grid,
                            double *out4, hipStream_t s) {
    if (dta == 0 && dtb == 0)
        hipLaunchKernelGGL((k_fid1<float, float>), dim3(grid), dim3(256), 0, s, (const C2<float> *)a, (const C2<float> *)b, n, part);
    else if (dta == 0)
        hipLaunchKernelGGL((k_fid1<float, double>), dim3(grid), dim3(256), 0, s, (const C2<float> *)a, (const C2<double> *)b, n, part);
    else if (dtb == 0)
        hipLaunchKernelGGL((k_fid1<double, float>), dim3(grid), dim3(256), 0, s, (const C2<double> *)a, (const C2<float> *)b, n, part);
    else
        hipLaunchKernelGGL((k_fid1<double, double>), dim3(grid), dim3(256), 0, s, (const C2<double> *)a, (const C2<double> *)b, n, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(256), 0, s, part, int64_t(grid), 4, out4, 0);
    return hipGetLastError();
}

hipError_t launch_fid_pass2(int dta, const void *a, int dtb, const void *b, int64_t n, const double *coef,
                            double *part, uint32_t grid, double *out2, hipStream_t s) {
    if (dta == 0 && dtb == 0)
        hipLaunchKernelGGL((k_fid2<float, float>), dim3(grid), dim3(256), 0, s, (const C2<float> *)a, (const C2<float> *)b, n, coef, part);
    else if (dta == 0)
        hipLaunchKernelGGL((k_fid2<float, double>), dim3(grid), dim3(256), 0, s, (const C2<float> *)a, (const C2<double> *)b, n, coef, part);
    else if (dtb == 0)
        hipLaunchKernelGGL((k_fid2<double, float>), dim3(grid), dim3(256), 0, s, (const C2<double> *)a, (const C2<float> *)b, n, coef, part);
    else
        hipLaunchKernelGGL((k_fid2<double, double>), dim3(grid), dim3(256), 0, s, (const C2<double> *)a, (const C2<double> *)b, n, coef, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(256), 0, s, part, int64_t(grid), 2, out2, 2);
    return hipGetLastError();
}

hipError_t set_kernel_attributes() {
    const int big = 160 * 1024;
    hipError_t e = hipSuccess;
#define SVX_ATTR(fn)                                                                              \
    if (e == hipSuccess) e = hipFuncSetAttribute(reinterpret_cast<const void *>(fn),               \
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, big);
    SVX_ATTR((k_pass<float, 10, false>)) SVX_ATTR((k_pass<float, 11, false>)) SVX_ATTR((k_pass<float, 12, false>))
    SVX_ATTR((k_pass<float, 13, false>)) SVX_ATTR((k_pass<float, 14, false>))
    SVX_ATTR((k_pass<float, 10, true>)) SVX_ATTR((k_pass<float, 11, true>)) SVX_ATTR((k_pass<float, 12, true>))
    SVX_ATTR((k_pass<float, 13, true>)) SVX_ATTR((k_pass<float, 14, true>))
    SVX_ATTR((k_pass_d4<12>)) SVX_ATTR((k_pass_d4<13>)) SVX_ATTR((k_pass_batch_d4<12>)) SVX_ATTR((k_pass_batch_d4<13>))
    SVX_ATTR((k_pass_batch<10, false>)) SVX_ATTR((k_pass_batch<10, true>)) SVX_ATTR((k_pass_batch<11, false>))
    SVX_ATTR((k_pass_batch<11, true>)) SVX_ATTR((k_pass_batch<12, false>)) SVX_ATTR((k_pass_batch<12, true>))
    SVX_ATTR((k_pass_batch<13, false>)) SVX_ATTR((k_pass_batch<13, true>)) SVX_ATTR((k_pass_batch<14, false>))
    SVX_ATTR((k_pass_batch<14, true>))
    SVX_ATTR((k_pass<double, 10, false>)) SVX_ATTR((k_pass<double, 11, false>)) SVX_ATTR((k_pass<double, 12, false>))
    SVX_ATTR((k_pass<double, 13, false>)) SVX_ATTR((k_permute<float>)) SVX_ATTR((k_permute<double>))
    if (e == hipSuccess) e = set_mux_attributes();
#undef SVX_ATTR
    return e;
}

}  // namespace svx
