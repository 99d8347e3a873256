// Internal descriptors shared by the host side (svx_api.hip) and the kernels (svx_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svx {

constexpr int kR = 5;            // register bits per group: each thread holds 2^5 amplitudes
constexpr int kNT = 256;         // threads per workgroup of the pass kernel (4 wavefronts)
constexpr int kNTD4 = 512;       // threads per workgroup of the dense-only pass kernel
constexpr int kMaxM = 8;         // gathered high qubits per pass (sub-block table 2^8 entries)
constexpr int kMinK = 10;        // smallest workgroup tile of the main pass kernel
constexpr int kMaxKc64 = 14;     // 2^14 complex64 = 128 KiB (+ padding) of LDS
constexpr int kMaxKc128 = 13;    // 2^13 complex128 = 128 KiB (+ padding)
constexpr int kSmallMaxBits = 9; // states of <= 2^9 amplitudes run in one workgroup (k_pass_small)
constexpr int kMaxTileSub = 8;   // at most 2^8 QSV2 tiles per workgroup tile

// op kinds (host classification of the MUX2 record; exact comparisons on the typed values)
enum OpKind : int32_t {
    K_FULL = 0,    // unconditional 2x2 M0 (cmask == 0 or M0 == M1)
    K_DIAG = 1,    // unconditional diagonal M0
    K_XSWAP = 2,   // unconditional exact X
    K_CFULL = 3,   // controlled, M0 == I: selected pairs get M1
    K_CDIAG = 4,   // controlled, M0 == I, M1 diagonal
    K_CPHASE = 5,  // controlled, M0 == I, M1 = diag(1, e)
    K_CXSWAP = 6,  // controlled, M0 == I, M1 exact X
    K_MUX = 7,     // controlled, general: selected pairs M1, others M0
    K_CXSWAP1 = 8, // controlled exact X, one control in the register group, none outside
    K_CPHASE1 = 9, // controlled diag(1, e), one control in the register group, none outside
    K_DENSE = 10,  // small-state kernel only: dense op, cmask = index of its GroupDev (type 2)
    K_TABLE = 11,  // complex64 pass kernel: diagonal table, 2^k complex64 phases at float offset aux
                   // of the dense operand array, indexed by the QSV2 tile index (no register bits)
    // op kinds of the FP32 MUX2 kernel (k_mux); the 16 floats m0, m1 hold the coefficients, and the
    // record's kind field holds (kind - M_FULL) * 8 + the target register bit
    M_FULL = 32,   // uncontrolled 2x2: w[4 i + j] = real 4x4 form, rows (y0r y0i y1r y1i), cols (x0r x0i x1r x1i)
    M_REAL = 33,   // uncontrolled real 2x2: w = (m00, m01, m10, m11)
    M_DIAG = 34,   // uncontrolled diagonal: w = (d0r, d0i, d1r, d1i)
    M_X = 35,      // uncontrolled exact X (a register exchange)
    M_CX = 36,     // controlled exact X (M0 = I): selected pairs exchange
    M_CPHASE = 37, // controlled diag(1, e) (M0 = I): w = (er, ei)
    M_CGEN = 38,   // any other controlled op: selected pairs get M1, the others M0 (complex64 m0, m1)
    M_TABLE = 39,  // diagonal table (aux: float offset of the slot-major table)
    M_FULLM = 40,  // M_FULL with half of the pairs on the FP32 matrix cores (v_mfma_f32_4x4x1_16b_f32)
    M_CX1 = 41,    // controlled exact X, one control among the register bits (aux: its register bit), none outside
    M_CXT = 42,    // controlled exact X, controls only outside the register bits (one lane mask for all pairs)
    M_CPHASE1 = 43,  // controlled diag(1, e), one control among the register bits (aux), none outside
    M_CPHASET = 44,  // controlled diag(1, e), controls only outside the register bits
};

struct PassDev {
    int32_t k, m, K, nbits;      // QSV2 tile bits, high count, workgroup tile bits, state bits
    int32_t ngroups, group0, nops, op0;
    int32_t h[kMaxM];            // physical bit of tile qubit k-m+i
    int32_t hs[kMaxM];           // h sorted ascending (tile-id deposit)
    int32_t tscale;              // > 0: scale each LDS tile so its largest component is < 2^tscale
                                 // (dense groups of type 4 rely on it); 0: no tile scaling
    int32_t torder;              // workgroup -> tile order: 0 identity, -1 bit reversal, -2 XCD-contiguous,
                                 // r > 0 rotate left by r
};

struct GroupDev {
    int32_t nops, op0;
    int32_t b[kR];               // workgroup-tile bit held by register bit q
    int32_t bs[kR];              // b sorted ascending (slot deposit)
    int32_t type;                // 0: MUX2 register group; 1: dense op, FP32 MFMA (32x32x2, complex as a
                                 // 2^(f+1) real matrix); 5: 4-qubit dense op, FP32 MFMA (16x16x4, three
                                 // real products); 2: dense op, small-state
                                 // kernel; 3: dense op, split-FP16 MFMA, per-column scaling; 4: the same
                                 // on a tile pre-scaled by the pass kernel (PassDev::tscale);
                                 // 6: diagonal table on its own (no MUX2 op beside it), 2^k complex64
                                 // phases at w_off indexed by the QSV2 tile index (complex64 k_pass)
    int32_t f;                   // dense: qubits of the op (type 1: 4 or 5 after padding)
    int32_t dq[5];               // dense: workgroup-tile bit of the op's qubit i (matrix index bit i)
    int32_t dqs[5];              // dq sorted ascending (column deposit)
    int32_t wexp;                // split-FP16 dense: the matrix was scaled by 2^wexp before the split
    int32_t exact;               // dense: entries all in {0, +-1, +-i} (FP32 products exact; statistics)
    int32_t jmask;               // type 5: the four tile bits that index the 16 columns of a column
                                 // block (lane l % 16); the block index runs over the other free bits
    int64_t w_off;               // dense: float offset of the matrix (type 1: MFMA A-operand order;
                                 // type 2: row-major re, im)
};

template <typename T>
struct OpDev {
    int32_t kind, tr;            // kind, target register bit
    uint32_t cmr, cmt;           // control mask over register bits / over other tile bits
    int32_t t, cmask;            // raw tile-coordinate target and control mask (small kernel)
    int32_t aux, pad_;           // K_TABLE: float offset of the table
    T m0[8], m1[8];              // (re, im) of m00 m01 m10 m11
};

struct PermDev {
    int32_t n, c, v, nfix;       // state bits, contiguous low bits, tile bits, fixed bits (n - v)
    int32_t fixnew[40];          // new bit position of fixed bit i (ascending)
    int32_t fixold[40];          // old bit position of the same bit
    uint16_t rdlo[64];           // LDS position (new-local index) of read-order index q, low 6 bits
    uint16_t rdhi[64];           // ... high part (q >> 6)
    uint64_t oldhi[64];          // old-index offset of read-order index q >> c
    uint64_t newhi[64];          // new-index offset of new-local index l >> c
};

// one state's pass in a batched launch (continuous batching)
struct BatchItem {
    void *st;                    // complex64 state
    const PassDev *pd;
    const GroupDev *gd;
    const void *od;              // OpDev<float> array of the prepared segment
    const float *wd;             // dense operand blocks of the prepared segment (or null)
    uint64_t tile0;              // first workgroup of this item in the launch
};

// persistent queue of whole-state circuits (k_mux_queue): records in host-visible memory
struct QueueJob {
    uint64_t st, pd, gd, od, wd;   // state, and the prepared pass's PassDev / GroupDev / OpDev / operands
    uint64_t flags;                // bit 0: start from |0...0>
};
struct QueueCtl {
    uint32_t tail;                 // jobs published (host: release store)
    uint32_t stop;                 // host asks every workgroup to leave
    uint32_t pad[14];
    uint32_t done[1];              // [capacity]: ticket + 1 once job `ticket` is complete (device: release)
};

// launchers (svx_kernels.hip)
hipError_t launch_pass_batch(int K, bool dense, bool d4, const BatchItem *items, int n_items, uint32_t grid,
                             size_t lds, hipStream_t s);
// dense-only passes of 4-qubit units (512 threads), K in {12, 13}
hipError_t launch_pass_d4(int K, void *st, const PassDev *pd, const GroupDev *gd, const OpDev<float> *od,
                          const float *wd, uint32_t grid, size_t lds, hipStream_t s);
hipError_t launch_pass_c64(int K, bool small, void *st, const PassDev *pd, const GroupDev *gd,
                           const OpDev<float> *od, const float *wd, uint32_t grid, size_t lds, hipStream_t s);
hipError_t launch_pass_c128(int K, bool small, void *st, const PassDev *pd, const GroupDev *gd,
                            const OpDev<double> *od, uint32_t grid, size_t lds, hipStream_t s);
hipError_t set_kernel_attributes();
// FP32 MUX2 kernel (complex64 passes of MUX2 ops and diagonal tables): 2^(K - R) threads, R register bits
#ifndef SVX_MUX_R13
#define SVX_MUX_R13 4
#endif
constexpr int kMuxR13 = SVX_MUX_R13;   // register bits at K = 13 (4: 512 threads, 5: 256 threads)
int mux_register_bits(int K);
hipError_t set_mux_attributes();
hipError_t launch_mux(int K, bool dense, void *st, const PassDev *pd, const GroupDev *gd, const OpDev<float> *od,
                      const float *wd, uint32_t grid, size_t lds, hipStream_t s);
hipError_t launch_mux_batch(int K, bool dense, const BatchItem *items, int n_items, uint32_t grid, size_t lds,
                            hipStream_t s);
hipError_t launch_mux_queue(int K, const QueueJob *jobs, QueueCtl *ctl, uint32_t *next, uint32_t cap,
                            uint64_t idle_ticks, uint32_t grid, size_t lds, hipStream_t s);
// double-buffered complex64 pass kernel (one QSV2 tile of 2^K per LDS buffer, K in {12, 13});
// register groups of 4 tile bits
size_t pass_lds_bytes(int K, int m, int k, int elem_bytes, bool small, int nbits);
hipError_t launch_permute(int dtype, void *dst, const void *src, const PermDev *pd, uint32_t grid,
                          size_t lds, hipStream_t s);
hipError_t launch_swap_words(void *a, void *b, int64_t offa, int64_t offb, int64_t nwords, hipStream_t s);
hipError_t launch_zero(int dtype, void *st, int64_t n_amps, hipStream_t s);
hipError_t launch_product(int dtype, void *st, int32_t n, const double *tlo, const double *thi,
                          int32_t nlo, hipStream_t s);
hipError_t launch_convert(void *dst, int dt_dst, const void *src, int dt_src, int64_t n, hipStream_t s);
// reductions: per-workgroup FP64 partials into `part` (device), folded by a second kernel
struct MargCfg {
    int32_t n, c, low8, nlow;    // state bits, chunk bits, thread-index bits used, marginal qubits
    int32_t nsa, nra, nsb, nrb, nshi, pad_;
    int32_t sa[8], sa_x[8];      // marginal bits inside the thread index (< low8) and their bin bits
    int32_t ra[8];               // the other thread-index bits
    int32_t sb[8], sb_x[8];      // marginal bits in [8, c) and their bin bits
    int32_t rb[8];               // non-marginal bits in [8, c)
    int32_t shi[8], shi_x[8];    // marginal bits >= c, as chunk-index bits, and their bin bits
};
hipError_t launch_marginals(int dtype, const void *st, const MargCfg *cfg_dev, int nbin, double *part,
                            uint32_t grid, double *out, hipStream_t s);
hipError_t launch_fid_pass1(int dta, const void *a, int dtb, const void *b, int64_t n_amps, double *part,
                            uint32_t grid, double *out4, hipStream_t s);
hipError_t launch_fid_pass2(int dta, const void *a, int dtb, const void *b, int64_t n_amps, const double *coef,
                            double *part, uint32_t grid, double *out2, hipStream_t s);

}  // namespace svx
