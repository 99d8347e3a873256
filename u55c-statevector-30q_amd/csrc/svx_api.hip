// libsvx C ABI (include/svx.h): stream validation, pass/group/op descriptor construction,
// device-resident descriptor caches, launches.
//
// A QSV2 segment (header [MAGIC, L, n_passes, k, 0,0,0,0]; per pass [m, h0, h1, n_ops, h2, h3, 0, 0];
// per op [OP_MUX2, t, cmask, 0 x 5, M0 8 x f32, M1 8 x f32], host/sv/isa.py:5-24) is parsed once
// into per-pass descriptors.  Each pass's ops are split into groups: maximal runs of consecutive
// ops whose targets span at most 5 tile bits (the register bits of the pass kernel); every op is
// classified (unconditional / controlled with M0 = I / general multiplexed; full / diagonal /
// phase / exact X) by exact comparison of its matrices, which only removes multiplications by
// exact zeros and ones.
#include <hip/hip_runtime.h>

#include "svx_half.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <functional>
#include <mutex>
#include <thread>
#include <chrono>
#include <set>
#include <string>
#include <vector>

#include "../../include/svx.h"
#include "svx_internal.h"

using namespace svx;

namespace {

constexpr uint32_t kMagic = 0x51535632u;
constexpr int kHdr = 8, kPassW = 8, kOpW = 24, kOpMux2 = 1, kOpHighX = 0x48;

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
int hipfail(hipError_t e, const char *what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return SVX_E_HIP;
}
#define HIPCHK(x)                                    \
    do {                                             \
        hipError_t e_ = (x);                         \
        if (e_ != hipSuccess) return hipfail(e_, #x); \
    } while (0)

// ------------------------------------------------------------------------------------------ options
// Normalised preparation options (the cache key of a prepared segment): every field explicit.
struct Opts {
    int32_t fuse, dense_math, kernel, diag, dense_only, tile_scale, tile_order, mux, width;
    float cost4, cost5;
};

void default_opts(svx_options *o) {
    std::memset(o, 0, sizeof(*o));
    o->size = (int32_t)sizeof(svx_options);
    o->fuse = SVX_FUSE_AUTO;
    o->dense_math = SVX_DENSE_F32;
    o->kernel = SVX_KERNEL_AUTO;
    o->diag_tables = 1;
    o->dense_only = 1;
    o->tile_scale = 1;
    o->tile_order = SVX_ORDER_TUNABLE;
    o->mux_engine = SVX_MUX_AUTO;
}

int normalize_opts(const svx_options *in, Opts *o) {
    svx_options d;
    default_opts(&d);
    if (in && in->size != (int32_t)sizeof(svx_options))
        return fail(SVX_E_ARG, "svx_options.size must be sizeof(svx_options)");
    const svx_options &s = in ? *in : d;
    std::memset(o, 0, sizeof(*o));
    if (s.fuse < SVX_FUSE_AUTO || s.fuse > SVX_FUSE_DENSE4) return fail(SVX_E_ARG, "bad svx_options.fuse");
    if (s.dense_math != SVX_DENSE_F32 && s.dense_math != SVX_DENSE_SPLIT16)
        return fail(SVX_E_ARG, "bad svx_options.dense_math");
    if (s.kernel < SVX_KERNEL_AUTO || s.kernel > SVX_KERNEL_SINGLE) return fail(SVX_E_ARG, "bad svx_options.kernel");
    if (s.tile_order < SVX_ORDER_TUNABLE || s.tile_order > SVX_ORDER_XCD)
        return fail(SVX_E_ARG, "bad svx_options.tile_order");
    if (s.mux_engine != SVX_MUX_AUTO && s.mux_engine != SVX_MUX_VALU) return fail(SVX_E_ARG, "bad svx_options.mux_engine");
    if (s.fuse_width < 0 || s.fuse_width > 5) return fail(SVX_E_ARG, "svx_options.fuse_width must be 0..5");
    o->fuse = s.fuse;
    o->dense_math = s.dense_math;
    o->kernel = s.kernel;
    o->diag = s.diag_tables ? 1 : 0;
    o->dense_only = s.dense_only ? 1 : 0;
    o->tile_scale = s.tile_scale ? 1 : 0;
    o->tile_order = s.tile_order;
    o->mux = s.mux_engine;
    o->width = s.fuse_width ? s.fuse_width : 4;
    // dense-unit costs in FP32 FMAs per amplitude (a general MUX2 op costs 8 on the vector units):
    // split-FP16 units run at the vector units' rate of about one general op each; an FP32 4-qubit
    // unit (64 matrix-core FMAs per amplitude) measured like about four
    const bool f32 = s.dense_math == SVX_DENSE_F32;
    o->cost4 = s.fuse_cost4 > 0 ? s.fuse_cost4 : (f32 ? 32.f : 8.f);
    o->cost5 = s.fuse_cost5 > 0 ? s.fuse_cost5 : (f32 ? 64.f : 16.f);
    if (s.fuse_cost4 < 0 || s.fuse_cost5 < 0) {   // explicit zero-cost request: fuse every block
        o->cost4 = std::max(0.f, s.fuse_cost4);
        o->cost5 = std::max(0.f, s.fuse_cost5);
    }
    return 0;
}

// ------------------------------------------------------------------------------------------ descriptor arena
// Prepared descriptors live in library-owned device slabs, carved into power-of-two blocks with one
// free list per size; a released descriptor's block is reused, never hipFree'd (hipFree would wait
// for the device).  Blocks above the slab size get their own allocation.
class Arena {
  public:
    static constexpr size_t kSlab = size_t(16) << 20;
    static constexpr int kMinLog = 12;   // 4 KiB
    int alloc(size_t bytes, void **out, size_t *got) {
        int lg = kMinLog;
        while ((size_t(1) << lg) < bytes) ++lg;
        const size_t sz = size_t(1) << lg;
        if (sz > kSlab) {
            HIPCHK(hipMalloc(out, sz));
            big_.insert(*out);
            held_ += sz;
            *got = sz;
            return 0;
        }
        auto &fl = free_[lg];
        if (fl.empty()) {
            void *slab = nullptr;
            HIPCHK(hipMalloc(&slab, kSlab));
            held_ += kSlab;
            for (size_t off = 0; off < kSlab; off += sz) fl.push_back(static_cast<unsigned char *>(slab) + off);
        }
        *out = fl.back();
        fl.pop_back();
        *got = sz;
        return 0;
    }
    void release(void *p, size_t sz) {
        if (big_.count(p)) {
            (void)hipFree(p);
            big_.erase(p);
            held_ -= sz;
            return;
        }
        int lg = kMinLog;
        while ((size_t(1) << lg) < sz) ++lg;
        free_[lg].push_back(p);
    }
    size_t held() const { return held_; }

  private:
    std::map<int, std::vector<void *>> free_;
    std::set<void *> big_;
    size_t held_ = 0;
};

// Scratch of svx_reduce_f64_noalloc: one copy per device, kept until the process ends.  Host code
// uses only the address looked up per device (DeviceState::scratch), never this name.
constexpr int kScratchMaxLow = 8;                                        // at most 2^8 bins
constexpr size_t kScratchPartBytes = size_t(2048) << kScratchMaxLow << 3;  // 2048 workgroups x 256 bins x FP64
constexpr size_t kScratchBytes = kScratchPartBytes + 4096 + (size_t(8) << kScratchMaxLow);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceState {
    bool ready = false;
    svx_caps caps{};
    unsigned char *scratch = nullptr;   // this device's address of the scratch array
    // library scratch for reductions and product-state tables
    double *part = nullptr;
    size_t part_bytes = 0;
    void *misc = nullptr;          // small descriptors / outputs
    size_t misc_bytes = 0;
    // batched launches: a ring of item tables (pinned host staging + device copy), one event per
    // slot so a slot is rewritten only after the launch that read it has finished
    static constexpr int kBatchSlots = 16;
    static constexpr int kBatchCap = 4096;
    BatchItem *batch_host = nullptr;
    BatchItem *batch_dev = nullptr;
    hipEvent_t batch_ev[kBatchSlots] = {};
    bool batch_used[kBatchSlots] = {};
    int batch_next = 0;
    std::mutex mu;
    Arena arena;                   // prepared descriptors (guarded by g_prep_mu)
};
std::mutex g_mu;
std::map<int, std::unique_ptr<DeviceState>> g_dev;

int current_device(int *dev) {
    HIPCHK(hipGetDevice(dev));
    return 0;
}

int ensure_device(int dev, DeviceState **out) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto &slot = g_dev[dev];
    if (!slot) slot.reset(new DeviceState());
    DeviceState *ds = slot.get();
    if (!ds->ready) {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, dev));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            return fail(SVX_E_NODEV, std::string("libsvx is built for gfx950, device is ") + prop.gcnArchName);
        HIPCHK(set_kernel_attributes());
        // this device's copy of the scratch array (the allocation-free reduction works in it)
        void *sp = nullptr;
        HIPCHK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
        ds->scratch = static_cast<unsigned char *>(sp);
        ds->caps.device = dev;
        ds->caps.cu_count = prop.multiProcessorCount;
        ds->caps.lds_per_block = (int64_t)prop.sharedMemPerBlock;
        ds->caps.hbm_bytes = (int64_t)prop.totalGlobalMem;
        ds->caps.wavefront = prop.warpSize;
        ds->caps.max_k_c64 = kMaxKc64;
        ds->caps.max_k_c128 = kMaxKc128;
        ds->caps.max_high = kMaxM;
        std::snprintf(ds->caps.arch, sizeof(ds->caps.arch), "%s", prop.gcnArchName);
        ds->ready = true;
    }
    *out = ds;
    return 0;
}

int get_ds(DeviceState **ds) {
    int dev = 0;
    int rc = current_device(&dev);
    if (rc) return rc;
    return ensure_device(dev, ds);
}

inline int ilog2_exact(int64_t x) {
    if (x <= 0 || (x & (x - 1))) return -1;
    int b = 0;
    while ((int64_t(1) << b) < x) ++b;
    return b;
}

// ------------------------------------------------------------------------------------------ parse
struct RawOp {
    int t;
    uint32_t cmask;
    float f[16];
    int dense_f = 0;           // fused-form record: number of qubits (0 for MUX2)
    int dq[5] = {0, 0, 0, 0, 0};
    std::vector<float> dm;     // 2^f x 2^f complex64, row-major (re, im)
};
constexpr int kOpDense = 2;
struct RawPass {
    int m;
    int h[kMaxM];
    std::vector<RawOp> ops;
};
struct RawSeg {
    int L, k;
    std::vector<RawPass> passes;
};

int parse_segment(const int32_t *w, int64_t nw, RawSeg *out) {
    if (nw < kHdr) return fail(SVX_E_FORMAT, "QSV2 stream shorter than its 8-word header");
    if ((uint32_t)w[0] != kMagic) return fail(SVX_E_MAGIC, "bad QSV2 magic");
    const int L = w[1], npass = w[2], k = w[3];
    if (L < 1 || L > 40) return fail(SVX_E_FORMAT, "QSV2 header: L out of range");
    if (k < 1 || k > L) return fail(SVX_E_FORMAT, "QSV2 header: tile bits k must satisfy 1 <= k <= L");
    if (npass < 0) return fail(SVX_E_FORMAT, "QSV2 header: negative pass count");
    out->L = L;
    out->k = k;
    out->passes.clear();
    int64_t pos = kHdr;
    for (int p = 0; p < npass; ++p) {
        if (pos + kPassW > nw) return fail(SVX_E_FORMAT, "QSV2 stream truncated in a pass header");
        RawPass rp;
        rp.m = w[pos];
        const int nops = w[pos + 3];
        int hv[4] = {w[pos + 1], w[pos + 2], w[pos + 4], w[pos + 5]};
        pos += kPassW;
        if (rp.m < 0 || rp.m > kMaxM || rp.m > k)
            return fail(SVX_E_FORMAT, "pass has m = " + std::to_string(rp.m) + " high qubits (max " +
                                          std::to_string(std::min(k, kMaxM)) + ")");
        for (int i = 0; i < std::min(rp.m, 4); ++i) rp.h[i] = hv[i];
        if (rp.m > 4) {
            if (pos + kPassW > nw || w[pos] != kOpHighX)
                return fail(SVX_E_FORMAT, "pass with m > 4 lacks its high-qubit record");
            for (int i = 4; i < rp.m; ++i) rp.h[i] = w[pos + 1 + (i - 4)];
            pos += kPassW;
        }
        const int lowb = k - rp.m;
        for (int i = 0; i < rp.m; ++i) {
            if (rp.h[i] < lowb || rp.h[i] >= L)
                return fail(SVX_E_FORMAT, "high qubit outside [k-m, L)");
            for (int j = 0; j < i; ++j)
                if (rp.h[j] == rp.h[i]) return fail(SVX_E_FORMAT, "repeated high qubit in a pass");
        }
        if (nops < 0) return fail(SVX_E_FORMAT, "negative op count");
        rp.ops.resize(nops);
        for (int o = 0; o < nops; ++o) {
            const int32_t *r = w + pos;
            if (pos + 8 > nw) return fail(SVX_E_FORMAT, "QSV2 stream truncated in an op record");
            if (r[0] == kOpDense) {
                // fused form: [OP_DENSE, f, q0..q4, 0] + 2 * 4^f float32 words
                const int f = r[1];
                if (f < 1 || f > 5) return fail(SVX_E_FORMAT, "dense record with f outside 1..5");
                const int64_t len = 8 + (int64_t(2) << (2 * f));
                if (pos + len > nw) return fail(SVX_E_FORMAT, "QSV2 stream truncated in a dense record");
                RawOp &ro = rp.ops[o];
                ro.dense_f = f;
                for (int i = 0; i < f; ++i) {
                    ro.dq[i] = r[2 + i];
                    if (ro.dq[i] < 0 || ro.dq[i] >= k) return fail(SVX_E_FORMAT, "dense op qubit outside the tile");
                    for (int i2 = 0; i2 < i; ++i2)
                        if (ro.dq[i2] == ro.dq[i]) return fail(SVX_E_FORMAT, "repeated dense op qubit");
                }
                ro.dm.resize(size_t(2) << (2 * f));
                std::memcpy(ro.dm.data(), r + 8, ro.dm.size() * sizeof(float));
                ro.t = ro.dq[0];
                ro.cmask = 0;
                pos += len;
                continue;
            }
            if (pos + kOpW > nw) return fail(SVX_E_FORMAT, "QSV2 stream truncated in an op record");
            if (r[0] != kOpMux2) return fail(SVX_E_FORMAT, "unknown opcode " + std::to_string(r[0]));
            if (r[1] < 0 || r[1] >= k) return fail(SVX_E_FORMAT, "op target outside the tile");
            if ((uint32_t)r[2] >> k) return fail(SVX_E_FORMAT, "op control mask outside the tile");
            rp.ops[o].t = r[1];
            rp.ops[o].cmask = (uint32_t)r[2];
            std::memcpy(rp.ops[o].f, r + 8, 16 * sizeof(float));
            pos += kOpW;
        }
        out->passes.push_back(std::move(rp));
    }
    return 0;
}

// ------------------------------------------------------------------------------------------ build
template <typename T>
bool is_id(const T *m) { return m[0] == T(1) && m[1] == 0 && m[2] == 0 && m[3] == 0 && m[4] == 0 && m[5] == 0 && m[6] == T(1) && m[7] == 0; }
template <typename T>
bool is_x(const T *m) { return m[0] == 0 && m[1] == 0 && m[2] == T(1) && m[3] == 0 && m[4] == T(1) && m[5] == 0 && m[6] == 0 && m[7] == 0; }
template <typename T>
bool is_diag(const T *m) { return m[2] == 0 && m[3] == 0 && m[4] == 0 && m[5] == 0; }
template <typename T>
bool same(const T *a, const T *b) {
    for (int e = 0; e < 8; ++e)
        if (!(a[e] == b[e])) return false;
    return true;
}

struct LaunchInfo {
    bool small;
    int K;
    uint64_t ntiles;
    uint32_t grid;
    uint32_t block;       // threads per workgroup
    size_t lds;
    int64_t pass_index;   // index into the PassDev array
    bool empty;           // no op changes the state (skip launch)
    bool dense;           // pass holds dense (fused-form) groups or tables: the dense instance of k_pass
    bool d4;              // every group a dense 4-qubit unit: the 512-thread dense-only kernel
    bool tunable;         // single-buffer / dense-only launch whose tile order svx_tune_prepared may switch
    bool mux;             // the FP32 MUX2 kernel k_mux (complex64 passes of MUX2 ops, diagonal tables and
                          // FP32 dense units)
    bool units;           // the pass holds dense units (not only tables): k_mux's dense instance
    int32_t stats[7];     // MUX2 groups, MUX2 ops, dense 4-qubit, dense 5-qubit, dense exact (FP32),
                          // MUX2 ops on the matrix cores, diagonal tables
};

struct Prepared {
    int dtype = 0, nbits = 0;
    std::vector<int32_t> words;
    Opts opts{};
    uint64_t key = 0;
    void *dev = nullptr;        // [PassDev x P][PassDev x P, identity tile order][GroupDev x G][OpDev<T> x O][operands]
    size_t dev_bytes = 0;       // arena block size
    size_t pass_off = 0, alt_off = 0, group_off = 0, op_off = 0, w_off = 0;
    std::vector<LaunchInfo> launches;
    // per pass: 1 = launch with the identity tile order (set by svx_tune_prepared, read by launches)
    std::unique_ptr<std::atomic<uint8_t>[]> alt;
    int torder = -2;            // default workgroup -> tile order of the passes (PassDev::torder)
    int device = 0;
    int refs = 0;               // references held by callers (svx_prepare / svx_release_prepared)
    uint64_t lru = 0;           // last release order (unreferenced entries are evicted oldest first)
    hipEvent_t last_use = nullptr;   // recorded after each launch sequence (not under stream capture)
};

// Registry of prepared segments.  handle = index + 1; an evicted handle's slot stays empty, so a
// stale handle is an error, never another segment.
std::mutex g_prep_mu;
std::vector<std::unique_ptr<Prepared>> g_prepared;
std::multimap<uint64_t, int64_t> g_prep_index;   // content hash -> handle
size_t g_unref_bytes = 0;                         // descriptor bytes of unreferenced entries
size_t g_unref_limit = size_t(256) << 20;
uint64_t g_lru_clock = 0;

Prepared *lookup_locked(int64_t handle) {
    if (handle < 1 || handle > (int64_t)g_prepared.size()) return nullptr;
    Prepared *p = g_prepared[handle - 1].get();
    return (p && p->refs > 0) ? p : nullptr;
}

uint64_t fnv1a(const int32_t *w, int64_t n, uint64_t seed) {
    uint64_t h = 1469598103934665603ull ^ seed;
    const unsigned char *p = reinterpret_cast<const unsigned char *>(w);
    for (int64_t i = 0; i < n * 4; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

// A dense unit: 2^f x 2^f complex matrix (row-major, re/im doubles) on tile bits q[0..f-1]; matrix
// index bit i <-> tile bit q[i].  From an OP_DENSE record or from pass fusion.
struct DenseUnit {
    int f = 0;
    int q[5] = {0, 0, 0, 0, 0};
    std::vector<double> u;
    // diagonal table (diag = true): the product of a run of diagonal ops over the tile bits dmask,
    // u = 2^k (re, im) entries indexed by the QSV2 tile index
    bool diag = false;
    uint32_t dmask = 0;
};

// The four tile bits that index the 16 columns of a 4-qubit unit's column block (lane l % 16 of a
// wavefront step).  A wavefront's 8-byte LDS access is served 16 lanes at a time, and the padded
// layout (element e at 8 (e + (e >> 5)) bytes) keeps those 16 lanes free of bank conflicts when their
// padded element offsets differ modulo 32 (64 banks of 4 bytes; measured on gfx950 over every 4-bit
// unit set and lane-bit choice of a 2^13 tile, this rule never passed a conflicting pattern).  The
// lowest four free bits conflict for some unit sets (for instance units on bits {3, 4, 5, 6}: columns
// on bits 0, 1, 2, 7); then the first conflict-free choice of four free bits is taken.
static uint32_t pad_elem(uint32_t e) { return e + (e >> 5); }
static bool lanes_conflict_free(uint32_t jmask) {
    uint32_t seen = 0;
    for (uint32_t j = 0; j < 16; ++j) {
        uint32_t e = 0, m = jmask;
        for (int i = 0; i < 4; ++i) {
            const int b = __builtin_ctz(m);
            m &= m - 1;
            e |= ((j >> i) & 1u) << b;
        }
        const uint32_t bank = pad_elem(e) & 31u;
        if ((seen >> bank) & 1u) return false;
        seen |= 1u << bank;
    }
    return true;
}
static uint32_t unit_lane_bits(int K, uint32_t unit_mask) {
    int fb[32], nf = 0;
    for (int b = 0; b < K; ++b)
        if (!((unit_mask >> b) & 1u)) fb[nf++] = b;
    const uint32_t lowest = (1u << fb[0]) | (1u << fb[1]) | (1u << fb[2]) | (1u << fb[3]);
    if (lanes_conflict_free(lowest)) return lowest;
    for (int a = 0; a < nf; ++a)
        for (int b = a + 1; b < nf; ++b)
            for (int c = b + 1; c < nf; ++c)
                for (int d = c + 1; d < nf; ++d) {
                    const uint32_t m = (1u << fb[a]) | (1u << fb[b]) | (1u << fb[c]) | (1u << fb[d]);
                    if (lanes_conflict_free(m)) return m;
                }
    return lowest;
}

// VALU cost of one op per amplitude, in FMAs (a general 2x2 MUX2 = 8), by OpKind
const double kOpCost[] = {8, 4, 1, 8, 4, 2, 2, 8, 2, 2};
// smallest MUX2 cost of a run of diagonal ops that becomes a table (one complex multiply and one
// table read per amplitude, plus an LDS round trip)
constexpr double kDiagRunCost = 8.0;

struct FuseCfg {
    bool all_dense = false; // every block (also single ops) becomes a dense unit
    int width = 4;          // largest block (tile bits): 4-qubit units run on the 512-thread kernel
    double cost4 = 8.0;     // cost of one dense op padded to 4 qubits, same units as kOpCost
    double cost5 = 16.0;    // ... 5 qubits
};

// Dense matrix of one MUX2 op on the block's tile bits qs (ascending; local index bit i <-> qs[i]):
// the pair (c0, c1) differing in the target bit gets M1 when the control mask is non-zero and all
// its bits are set in c0 (host/sv/model.py:59-71), else M0.
template <typename T>
void mux_matrix(const OpDev<T> &op, const std::vector<int> &qs, std::vector<double> &m) {
    const int f = (int)qs.size(), D = 1 << f;
    m.assign(size_t(2) * D * D, 0.0);
    int tp = 0;
    for (int i = 0; i < f; ++i)
        if (qs[i] == op.t) tp = i;
    const uint32_t cm = (uint32_t)op.cmask;
    for (int c = 0; c < D; ++c) {
        const int tb = (c >> tp) & 1;
        const int c0 = c & ~(1 << tp), c1 = c | (1 << tp);
        uint32_t tile0 = 0;
        for (int i = 0; i < f; ++i)
            if ((c0 >> i) & 1) tile0 |= 1u << qs[i];
        const T *M = (cm != 0 && (tile0 & cm) == cm) ? op.m1 : op.m0;
        // column c: out[c0] += M[0][tb] x_c, out[c1] += M[1][tb] x_c
        m[2 * (size_t(c0) * D + c)] = (double)M[2 * tb];
        m[2 * (size_t(c0) * D + c) + 1] = (double)M[2 * tb + 1];
        m[2 * (size_t(c1) * D + c)] = (double)M[4 + 2 * tb];
        m[2 * (size_t(c1) * D + c) + 1] = (double)M[4 + 2 * tb + 1];
    }
}

// a <- b a (complex D x D, row-major re/im)
void cmatmul_left(const std::vector<double> &b, std::vector<double> &a, int D) {
    std::vector<double> r(size_t(2) * D * D, 0.0);
    for (int i = 0; i < D; ++i)
        for (int l = 0; l < D; ++l) {
            const double br = b[2 * (size_t(i) * D + l)], bi = b[2 * (size_t(i) * D + l) + 1];
            if (br == 0.0 && bi == 0.0) continue;
            for (int j = 0; j < D; ++j) {
                const double ar = a[2 * (size_t(l) * D + j)], ai = a[2 * (size_t(l) * D + j) + 1];
                r[2 * (size_t(i) * D + j)] += br * ar - bi * ai;
                r[2 * (size_t(i) * D + j) + 1] += br * ai + bi * ar;
            }
        }
    a.swap(r);
}

// Pass fusion (complex64): the pass's ops are clustered over their dependency DAG (two ops depend
// when their target/control bits intersect) into blocks of at most cfg.width tile bits, each block
// grown by the ready op that adds the fewest new bits (earliest on ties); a block whose MUX2 cost
// exceeds the cost of one dense op becomes one dense unit (product of its ops in FP64), the others
// keep their ops.  A small beam search over the block sequence (other seed ops for each block) can
// replace that greedy partition when it needs fewer dense units.  Only ops on disjoint bits change
// order, so the result equals the pass's product of ops up to rounding.
template <typename T>
void fuse_pass(std::vector<OpDev<T>> &pops, std::vector<int> &dense_of, std::vector<DenseUnit> &dus,
               const FuseCfg &cfg) {
    const int n = (int)pops.size();
    std::vector<uint32_t> qm(n);
    std::vector<char> fusable(n);
    for (int i = 0; i < n; ++i) {
        if (dense_of[i] >= 0) {
            const DenseUnit &du = dus[dense_of[i]];
            uint32_t m = du.dmask;
            for (int q = 0; q < du.f; ++q) m |= 1u << du.q[q];
            qm[i] = m;
            fusable[i] = 0;
        } else {
            qm[i] = (1u << pops[i].t) | (uint32_t)pops[i].cmask;
            fusable[i] = __builtin_popcount(qm[i]) <= cfg.width;
        }
    }
    std::vector<int> npred(n, 0);
    std::vector<std::vector<int>> succ(n);
    std::vector<uint64_t> pm(n, 0);   // predecessors as a bit mask (passes of at most 64 ops)
    int last[32];
    for (int b = 0; b < 32; ++b) last[b] = -1;
    for (int i = 0; i < n; ++i) {
        std::vector<int> deps;
        for (int b = 0; b < 32; ++b)
            if ((qm[i] >> b) & 1u) {
                if (last[b] >= 0 && std::find(deps.begin(), deps.end(), last[b]) == deps.end()) deps.push_back(last[b]);
                last[b] = i;
            }
        npred[i] = (int)deps.size();
        for (int d : deps) {
            succ[d].push_back(i);
            if (d < 64) pm[i] |= uint64_t(1) << d;
        }
    }
    struct Block {
        std::vector<int> mem;
        bool single = false;   // a non-fusable op (wide controls, a dense record or table) alone
    };
    // greedy partition: each block grown from the ready ops by the one adding the fewest new bits
    std::vector<Block> greedy;
    {
        std::vector<int> np = npred;
        std::set<int> ready;
        for (int i = 0; i < n; ++i)
            if (np[i] == 0) ready.insert(i);
        while (!ready.empty()) {
            uint32_t S = 0;
            Block blk;
            while (true) {
                int best = -1, bestadd = 99;
                for (int i : ready) {
                    if (!fusable[i]) continue;
                    const int add = __builtin_popcount(qm[i] & ~S);
                    if (__builtin_popcount(S) + add > cfg.width) continue;
                    if (add < bestadd) {
                        best = i;
                        bestadd = add;
                    }
                }
                if (best < 0) {
                    if (!blk.mem.empty()) break;
                    best = *ready.begin();
                    blk.single = true;
                }
                ready.erase(best);
                blk.mem.push_back(best);
                S |= qm[best];
                for (int s2 : succ[best])
                    if (--np[s2] == 0) ready.insert(s2);
                if (blk.single) break;
            }
            greedy.push_back(std::move(blk));
        }
    }
    auto is_dense = [&](const Block &b) {
        if (b.single) return false;
        double mux = 0;
        uint32_t S = 0;
        for (int i : b.mem) {
            mux += kOpCost[pops[i].kind];
            S |= qm[i];
        }
        const double dcost = std::max(4, __builtin_popcount(S)) == 5 ? cfg.cost5 : cfg.cost4;
        return cfg.all_dense || (b.mem.size() >= 2 && mux > dcost);
    };
    auto score = [&](const std::vector<Block> &bl, int &units, int &left) {
        units = left = 0;
        for (const Block &b : bl) {
            if (is_dense(b)) ++units;
            else left += (int)b.mem.size();
        }
    };
    std::vector<Block> chosen = std::move(greedy);
    // beam search over block sequences (each block seeded by one ready op and grown like the greedy
    // one), keeping the 8 partial partitions that have placed the most ops; its partition replaces
    // the greedy one when it needs fewer dense units and leaves no more MUX2 ops
    if (n <= 64 && !cfg.all_dense) {
        struct St {
            uint64_t done;
            std::vector<Block> path;
        };
        const uint64_t full = n == 64 ? ~uint64_t(0) : ((uint64_t(1) << n) - 1);
        std::vector<St> cur(1, St{0, {}});
        std::vector<Block> found;
        for (int depth = 0; depth < n && found.empty() && !cur.empty(); ++depth) {
            std::vector<St> nxt;
            std::set<uint64_t> seen;
            for (const St &st : cur) {
                for (int r = 0; r < n; ++r) {
                    if (((st.done >> r) & 1u) || (pm[r] & ~st.done)) continue;
                    uint64_t d = st.done | (uint64_t(1) << r);
                    Block b;
                    b.mem.push_back(r);
                    if (!fusable[r]) {
                        b.single = true;
                    } else {
                        uint32_t S = qm[r];
                        while (true) {
                            int best = -1, bestadd = 99;
                            for (int i = 0; i < n; ++i) {
                                if (((d >> i) & 1u) || !fusable[i] || (pm[i] & ~d)) continue;
                                const int add = __builtin_popcount(qm[i] & ~S);
                                if (__builtin_popcount(S) + add > cfg.width) continue;
                                if (add < bestadd) {
                                    best = i;
                                    bestadd = add;
                                }
                            }
                            if (best < 0) break;
                            d |= uint64_t(1) << best;
                            b.mem.push_back(best);
                            S |= qm[best];
                        }
                    }
                    if (!seen.insert(d).second) continue;
                    St ns{d, st.path};
                    ns.path.push_back(std::move(b));
                    if (d == full && found.empty()) found = ns.path;
                    nxt.push_back(std::move(ns));
                }
            }
            std::stable_sort(nxt.begin(), nxt.end(), [](const St &a, const St &b) {
                return __builtin_popcountll(a.done) > __builtin_popcountll(b.done);
            });
            if (nxt.size() > 8) nxt.resize(8);
            cur.swap(nxt);
        }
        if (!found.empty()) {
            int gu, gl, bu, bl;
            score(chosen, gu, gl);
            score(found, bu, bl);
            if (bu < gu && bl <= gl) chosen = std::move(found);
        }
    }
    std::vector<OpDev<T>> out;
    std::vector<int> out_dense;
    for (const Block &blk : chosen) {
        if (is_dense(blk)) {
            uint32_t S = 0;
            for (int i : blk.mem) S |= qm[i];
            const int nq = __builtin_popcount(S);
            std::vector<int> qs;
            for (int b = 0; b < 32; ++b)
                if ((S >> b) & 1u) qs.push_back(b);
            const int D = 1 << nq;
            DenseUnit du;
            du.f = nq;
            for (int q = 0; q < nq; ++q) du.q[q] = qs[q];
            du.u.assign(size_t(2) * D * D, 0.0);
            for (int dd = 0; dd < D; ++dd) du.u[2 * (size_t(dd) * D + dd)] = 1.0;
            std::vector<double> m;
            for (int i : blk.mem) {
                mux_matrix(pops[i], qs, m);
                cmatmul_left(m, du.u, D);
            }
            OpDev<T> od{};
            od.kind = -1;
            od.t = qs[0];
            out.push_back(od);
            out_dense.push_back((int)dus.size());
            dus.push_back(std::move(du));
        } else {
            for (int i : blk.mem) {
                out.push_back(pops[i]);
                out_dense.push_back(dense_of[i]);
            }
        }
    }
    pops.swap(out);
    dense_of.swap(out_dense);
}

// Diagonal tables (complex64): diagonal ops commute with each other, so a run of them -- hoisted
// over the ops between them that share no bit with them -- is one diagonal operator over the QSV2
// tile index, applied as a table of 2^k phases (one complex multiply per amplitude instead of the
// run's MUX2 work).  Scanning in pass order, a diagonal op joins the open run unless it shares a bit
// with an op already deferred behind the run; other ops are deferred (kept in order, after the
// run).  A run of at least two ops whose MUX2 cost reaches min_cost becomes a table, computed in
// FP64 from the ops' complex64 entries with the reference's selection rule (pair index i0 with the
// target bit clear, M1 iff cmask != 0 and (i0 & cmask) == cmask: host/sv/model.py:43-56).
template <typename T>
void diag_runs(std::vector<OpDev<T>> &pops, std::vector<int> &dense_of, std::vector<DenseUnit> &dus, int k,
               double min_cost) {
    auto bits = [&](size_t i) -> uint32_t {
        if (dense_of[i] < 0) return (1u << pops[i].t) | (uint32_t)pops[i].cmask;
        const DenseUnit &du = dus[dense_of[i]];
        uint32_t m = du.dmask;
        for (int q = 0; q < du.f; ++q) m |= 1u << du.q[q];
        return m;
    };
    auto diag = [&](size_t i) {
        return dense_of[i] < 0 && is_diag(pops[i].m0) && (pops[i].cmask == 0 || is_diag(pops[i].m1));
    };
    std::vector<OpDev<T>> out;
    std::vector<int> out_dense;
    std::vector<size_t> run, def;
    uint32_t mrun = 0, mdef = 0;
    double rcost = 0;
    auto flush = [&]() {
        if (run.size() >= 2 && rcost >= min_cost) {
            const size_t D = size_t(1) << k;
            DenseUnit du;
            du.diag = true;
            du.dmask = mrun;
            du.u.assign(2 * D, 0.0);
            for (size_t x = 0; x < D; ++x) {
                double re = 1.0, im = 0.0;
                for (size_t i : run) {
                    const OpDev<T> &o = pops[i];
                    const uint32_t i0 = (uint32_t)x & ~(1u << o.t);
                    const bool sel = o.cmask != 0 && (i0 & (uint32_t)o.cmask) == (uint32_t)o.cmask;
                    const T *M = sel ? o.m1 : o.m0;
                    const int e = ((x >> o.t) & 1) ? 6 : 0;   // m11 or m00
                    const double dr = (double)M[e], di = (double)M[e + 1];
                    const double nr = re * dr - im * di;
                    im = re * di + im * dr;
                    re = nr;
                }
                du.u[2 * x] = re;
                du.u[2 * x + 1] = im;
            }
            OpDev<T> od{};
            od.kind = -1;
            od.t = __builtin_ctz(mrun);
            out.push_back(od);
            out_dense.push_back((int)dus.size());
            dus.push_back(std::move(du));
        } else {
            for (size_t i : run) {
                out.push_back(pops[i]);
                out_dense.push_back(-1);
            }
        }
        for (size_t i : def) {
            out.push_back(pops[i]);
            out_dense.push_back(dense_of[i]);
        }
        run.clear();
        def.clear();
        mrun = mdef = 0;
        rcost = 0;
    };
    for (size_t i = 0; i < pops.size(); ++i) {
        const uint32_t b = bits(i);
        if (diag(i)) {
            if (b & mdef) flush();
            run.push_back(i);
            mrun |= b;
            rcost += kOpCost[pops[i].kind];
        } else if (run.empty()) {
            out.push_back(pops[i]);
            out_dense.push_back(dense_of[i]);
        } else {
            def.push_back(i);
            mdef |= b;
        }
    }
    flush();
    pops.swap(out);
    dense_of.swap(out_dense);
}

// Register groups of the FP32 MUX2 kernel.  Ops are list-scheduled over their commutation DAG:
// two ops commute when neither one's mixing bit (the target of a non-diagonal op; none for a
// diagonal op or table) lies in the other's support (target, controls, table bits), so only
// commuting ops change order and the pass's product of ops is unchanged.  A group collects ready
// ops while their register targets fit R bits, preferring ops whose target is already a register
// bit; a controlled phase diag(1, e) acts symmetrically on all its bits, so any of them may be its
// register target (the others become controls).  Each op becomes an M_* record with its
// coefficients in the form the kernel consumes.
template <typename T, typename PutTable>
int mux_groups(const std::vector<OpDev<T>> &pops, const std::vector<int> &dense_of, const std::vector<DenseUnit> &dus,
               int K, int R, bool mfma, std::vector<GroupDev> &groups, std::vector<OpDev<T>> &ops,
               PutTable &&put_table_regs, double &log2_growth, int32_t *n_mfma) {
    const int n = (int)pops.size();
    std::vector<uint32_t> sup(n), mix(n), tc(n);
    for (int i = 0; i < n; ++i) {
        if (dense_of[i] >= 0) {   // diagonal table
            sup[i] = dus[dense_of[i]].dmask;
            mix[i] = 0;
            tc[i] = 0;
            continue;
        }
        const OpDev<T> &o = pops[i];
        sup[i] = (1u << o.t) | (uint32_t)o.cmask;
        const bool diagonal = is_diag(o.m0) && (o.cmask == 0 || is_diag(o.m1));
        mix[i] = diagonal ? 0u : (1u << o.t);
        const bool cphase = o.kind == K_CPHASE || o.kind == K_CPHASE1;
        tc[i] = cphase ? sup[i] : (1u << o.t);
    }
    // dependency DAG (any number of ops: a whole-state pass may hold hundreds): op j waits for each
    // earlier op it does not commute with
    std::vector<std::vector<int>> succ(n);
    std::vector<int> waiting(n, 0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i)
            if ((mix[i] & sup[j]) || (mix[j] & sup[i])) {
                succ[i].push_back(j);
                ++waiting[j];
            }
    std::vector<char> placed(n, 0);
    int n_placed = 0;
    while (n_placed < n) {
        uint32_t S = 0;
        int nS = 0;
        std::vector<std::pair<int, int>> mem;   // (op, register target tile bit; -1 for a table)
        while (true) {
            int best = -1, bestcost = 99, bestbit = -1;
            for (int i = 0; i < n; ++i) {
                if (placed[i] || waiting[i]) continue;
                int cost, bit;
                if (tc[i] == 0) {
                    cost = 0;
                    bit = -1;
                } else if (tc[i] & S) {
                    cost = 0;
                    bit = __builtin_ctz(tc[i] & S);
                    if ((1u << pops[i].t) & S) bit = pops[i].t;   // keep the record's own target when possible
                } else if (nS < R) {
                    cost = 1;
                    bit = pops[i].t;
                } else {
                    continue;
                }
                if (cost < bestcost) {
                    best = i;
                    bestcost = cost;
                    bestbit = bit;
                }
            }
            if (best < 0) break;
            placed[best] = 1;
            ++n_placed;
            for (int j : succ[best]) --waiting[j];
            if (bestbit >= 0 && !((S >> bestbit) & 1u)) {
                S |= 1u << bestbit;
                ++nS;
            }
            mem.push_back({best, bestbit});
        }
        if (mem.empty()) return fail(SVX_E_FORMAT, "internal: register grouping made no progress");
        std::vector<int> rb;
        for (int b = 0; b < K; ++b)
            if ((S >> b) & 1u) rb.push_back(b);
        for (int b = K - 1; b >= 0 && (int)rb.size() < R; --b)
            if (!((S >> b) & 1u)) rb.push_back(b);
        GroupDev gd{};
        gd.nops = (int)mem.size();
        gd.op0 = (int)ops.size();
        uint32_t regmask = 0;
        for (int q = 0; q < R; ++q) {
            gd.b[q] = gd.bs[q] = rb[q];
            regmask |= 1u << rb[q];
        }
        std::sort(gd.bs, gd.bs + R);
        for (const auto &mb : mem) {
            const int i = mb.first;
            OpDev<T> od{};
            if (dense_of[i] >= 0) {
                od.kind = (M_TABLE - M_FULL) * 8;
                od.aux = (int32_t)put_table_regs(dus[dense_of[i]], rb);
                ops.push_back(od);
                continue;
            }
            const OpDev<T> &src = pops[i];
            const int t = mb.second;
            uint32_t cmask = (uint32_t)src.cmask;
            if (t != src.t) cmask = (cmask | (1u << src.t)) & ~(1u << t);   // controlled phase on another of its bits
            od.t = t;
            od.cmask = (int32_t)cmask;
            od.tr = (int)(std::find(rb.begin(), rb.end(), t) - rb.begin());
            const bool unc = src.kind <= K_XSWAP;
            uint32_t cmr = 0;
            for (int q = 0; q < R; ++q)
                if ((cmask >> rb[q]) & 1u) cmr |= 1u << q;
            od.cmr = unc ? 0u : cmr;
            od.cmt = unc ? 0u : (cmask & ~regmask);
            const T *m0 = src.m0, *m1 = src.m1;
            switch (src.kind) {
            case K_FULL: {
                const bool realm = m0[1] == 0 && m0[3] == 0 && m0[5] == 0 && m0[7] == 0;
                if (realm) {
                    od.kind = M_REAL;
                    od.m0[0] = m0[0];
                    od.m0[1] = m0[2];
                    od.m0[2] = m0[4];
                    od.m0[3] = m0[6];
                } else {
                    // general ops on the matrix cores (measured: 244.5 ms per unfused 30-qubit
                    // step, against 246.5 alternating with the vector units and 254 on the vector
                    // units alone)
                    od.kind = mfma ? M_FULLM : M_FULL;
                    T *w = od.m0;   // 16 floats over m0, m1
                    for (int r = 0; r < 2; ++r) {
                        const T ar = m0[4 * r], ai = m0[4 * r + 1], br = m0[4 * r + 2], bi = m0[4 * r + 3];
                        T *re = w + 8 * r, *im = w + 8 * r + 4;
                        re[0] = ar; re[1] = -ai; re[2] = br; re[3] = -bi;
                        im[0] = ai; im[1] = ar; im[2] = bi; im[3] = br;
                    }
                }
                break;
            }
            case K_DIAG:
                od.kind = M_DIAG;
                od.m0[0] = m0[0];
                od.m0[1] = m0[1];
                od.m0[2] = m0[6];
                od.m0[3] = m0[7];
                break;
            case K_XSWAP: od.kind = M_X; break;
            case K_CXSWAP:
            case K_CXSWAP1:
            case K_CPHASE:
            case K_CPHASE1: {
                // one register control and none outside: the touched pairs are fixed (CB in the cmr
                // slot); controls only outside: one lane mask; otherwise the general selection
                const bool phase = src.kind == K_CPHASE || src.kind == K_CPHASE1;
                if (od.cmt == 0 && __builtin_popcount(od.cmr) == 1) {
                    od.kind = phase ? M_CPHASE1 : M_CX1;
                    od.cmr = (uint32_t)__builtin_ctz(od.cmr);
                } else if (od.cmr == 0) {
                    od.kind = phase ? M_CPHASET : M_CXT;
                } else {
                    od.kind = phase ? M_CPHASE : M_CX;
                }
                if (phase) {
                    od.m0[0] = m1[6];
                    od.m0[1] = m1[7];
                }
                break;
            }
            default:
                od.kind = M_CGEN;
                for (int e = 0; e < 8; ++e) {
                    od.m0[e] = m0[e];
                    od.m1[e] = m1[e];
                }
                break;
            }
            auto rows = [](const T *mm) {
                const double a = std::hypot((double)mm[0], (double)mm[1]) + std::hypot((double)mm[2], (double)mm[3]);
                const double b = std::hypot((double)mm[4], (double)mm[5]) + std::hypot((double)mm[6], (double)mm[7]);
                return std::max(a, b);
            };
            log2_growth += std::log2(std::max(1.0, std::max(rows(m0), rows(m1))));
            if (od.kind == M_FULLM) ++*n_mfma;
            od.kind = (od.kind - M_FULL) * 8 + od.tr;   // the kernel dispatches on kind and target at once
            ops.push_back(od);
        }
        groups.push_back(gd);
    }
    return 0;
}

template <typename T>
int build_prepared(const RawSeg &seg, int nbits, int dtype, const Opts &opt, Prepared *pr,
                   std::vector<unsigned char> *blob) {
    const int k = seg.k;
    const int maxK = dtype == SVX_C64 ? kMaxKc64 : kMaxKc128;
    if (k > maxK)
        return fail(SVX_E_CAPACITY, "tile of 2^" + std::to_string(k) + " amplitudes exceeds the LDS tile limit 2^" +
                                        std::to_string(maxK));
    const bool small = nbits <= kSmallMaxBits;
    // kernel choice per pass (SVX_KERNEL_AUTO): dense-only passes of 4-qubit units -> k_pass_d4,
    // everything else -> the single-buffer kernels (k_mux for complex64 MUX2 / FP32 passes, k_pass
    // otherwise: two workgroups per CU, non-temporal 16-byte HBM access, register groups, dense
    // groups on the matrix cores).
    const bool dense_f32 = opt.dense_math == SVX_DENSE_F32;
    const bool tile_scale_ok = opt.tile_scale != 0;
    FuseCfg fcfg;
    fcfg.width = std::min(5, std::max(1, opt.width));
    fcfg.cost4 = opt.cost4;
    fcfg.cost5 = opt.cost5;
    if (opt.fuse == SVX_FUSE_DENSE4) {   // every block of at most 4 tile bits a dense unit
        fcfg.width = 4;
        fcfg.all_dense = true;
    }
    const int torder = opt.tile_order == SVX_ORDER_IDENTITY ? 0 : -2;   // default: each XCD a contiguous range
    const bool diag_ok = opt.diag != 0;
    int Kv1 = 0;
    if (!small) {
        Kv1 = std::max(k, kMinK);
        if (Kv1 - k > kMaxTileSub) Kv1 = k + kMaxTileSub;
        if (Kv1 < kMinK) Kv1 = kMinK;
        if (Kv1 > nbits) Kv1 = nbits;
        if (Kv1 > maxK) Kv1 = maxK;
        if (Kv1 < k) return fail(SVX_E_CAPACITY, "tile larger than the state");
    }
    DeviceState *ds = nullptr;
    int rc0 = get_ds(&ds);
    if (rc0) return rc0;
    std::vector<PassDev> passes;
    std::vector<GroupDev> groups;
    std::vector<OpDev<T>> ops;
    std::vector<LaunchInfo> launches;
    std::vector<float> wmat;      // dense matrices, MFMA A-operand order
    for (const RawPass &rp : seg.passes) {
        // classify
        std::vector<OpDev<T>> pops;
        std::vector<DenseUnit> dus;
        std::vector<int> dense_of;   // parallel to pops: index into dus, or -1 for a MUX2 op
        double cost = 0;
        bool has_dense = false;
        for (const RawOp &ro : rp.ops) {
            if (ro.dense_f) {
                if (dtype != SVX_C64)
                    return fail(SVX_E_ARG, "dense (fused-form) records run in complex64; complex128 runs use the plain MUX2 records");
                OpDev<T> od{};
                od.kind = -1;
                od.t = ro.dq[0];
                pops.push_back(od);
                DenseUnit du;
                du.f = ro.dense_f;
                for (int q = 0; q < du.f; ++q) du.q[q] = ro.dq[q];
                du.u.assign(ro.dm.begin(), ro.dm.end());
                dense_of.push_back((int)dus.size());
                dus.push_back(std::move(du));
                has_dense = true;
                cost += 8.0 * (ro.dense_f <= 4 ? 16 : 32) / 4.0;
                continue;
            }
            dense_of.push_back(-1);
            OpDev<T> od{};
            for (int e = 0; e < 8; ++e) {
                od.m0[e] = T(ro.f[e]);
                od.m1[e] = T(ro.f[8 + e]);
            }
            od.t = ro.t;
            od.cmask = (int32_t)ro.cmask;
            const bool unc = ro.cmask == 0 || same(od.m0, od.m1);
            if (small) {
                od.kind = K_MUX;
                pops.push_back(od);
                continue;
            }
            if (unc) {
                if (is_id(od.m0)) {   // exact no-op
                    dense_of.pop_back();
                    continue;
                }
                od.kind = is_x(od.m0) ? K_XSWAP : (is_diag(od.m0) ? K_DIAG : K_FULL);
            } else if (is_id(od.m0)) {
                if (is_x(od.m1)) od.kind = K_CXSWAP;
                else if (is_diag(od.m1) && od.m1[0] == T(1) && od.m1[1] == T(0)) od.kind = K_CPHASE;
                else if (is_diag(od.m1)) od.kind = K_CDIAG;
                else od.kind = K_CFULL;
            } else {
                od.kind = K_MUX;
            }
            cost += kOpCost[od.kind];
            pops.push_back(od);
        }
        if (dtype == SVX_C64 && !small && diag_ok && opt.fuse != SVX_FUSE_DENSE4) {
            diag_runs(pops, dense_of, dus, k, kDiagRunCost);
            for (int d : dense_of) has_dense = has_dense || d >= 0;
        }
        if (dtype == SVX_C64 && !small && opt.fuse != SVX_FUSE_OFF) {
            fuse_pass(pops, dense_of, dus, fcfg);
            for (int d : dense_of) has_dense = has_dense || d >= 0;
        }
        const int mode = small ? 0 : 1;   // 0 small-state kernel, 1 single-buffer kernels
        const int K = mode == 0 ? nbits : Kv1;
        // complex64 passes without dense units (diagonal tables allowed) run on the FP32 MUX2 kernel
        bool has_units = false;
        for (size_t i = 0; i < dense_of.size(); ++i) has_units = has_units || (dense_of[i] >= 0 && !dus[dense_of[i]].diag);
        // the MUX2 kernel: complex64 passes of MUX2 ops and tables, and of FP32 dense units too (the
        // split-FP16 units need the tile scaling of k_pass)
        const bool use_mux = mode == 1 && dtype == SVX_C64 && (!has_units || dense_f32);
        const int R = use_mux ? mux_register_bits(K) : kR;
        PassDev pd{};
        pd.k = k;
        pd.m = rp.m;
        pd.K = K;
        pd.nbits = nbits;
        pd.torder = torder;
        for (int i = 0; i < rp.m; ++i) pd.h[i] = pd.hs[i] = rp.h[i];
        std::sort(pd.hs, pd.hs + rp.m);
        pd.group0 = (int)groups.size();
        pd.op0 = (int)ops.size();
        pd.nops = (int)pops.size();
        LaunchInfo li{};
        li.small = mode == 0;
        li.dense = has_dense;
        li.K = K;
        li.pass_index = (int64_t)passes.size();
        li.empty = pd.nops == 0;
        li.ntiles = uint64_t(1) << (nbits - K);
        // the tile order matters for HBM access once many workgroups run at a time; a fixed order
        // option pins it
        li.tunable = mode == 1 && opt.tile_order == SVX_ORDER_TUNABLE && li.ntiles >= 64;
        if (mode == 0) {
            li.grid = 1;
            li.block = kNT;
            li.lds = 0;
        } else {
            // one workgroup per tile: the grid in work-items must stay below 2^32 (HIP's limit)
            if (li.ntiles * kNTD4 > 0xffffffffull)
                return fail(SVX_E_CAPACITY, "2^" + std::to_string(nbits - K) + " workgroup tiles of 2^" + std::to_string(K) +
                                                " amplitudes exceed the launch limit of 2^32 work-items");
            li.grid = (uint32_t)li.ntiles;
            li.block = use_mux ? (1u << (K - R)) : kNT;
            li.mux = use_mux;
            li.units = use_mux && has_units;
            li.lds = pass_lds_bytes(K, rp.m, k, (int)sizeof(T) * 2, false, nbits);
        }
        launches.push_back(li);
        if (mode == 0) {
            for (size_t i = 0; i < pops.size(); ++i) {
                OpDev<T> o = pops[i];
                if (dense_of[i] >= 0) {   // raw matrix for the small-state kernel's direct product
                    const DenseUnit &du = dus[dense_of[i]];
                    GroupDev gd{};
                    gd.type = 2;
                    gd.f = du.f;
                    for (int q = 0; q < du.f; ++q) gd.dq[q] = gd.dqs[q] = du.q[q];
                    gd.w_off = (int64_t)wmat.size();
                    for (double v : du.u) wmat.push_back((float)v);
                    o.kind = K_DENSE;
                    o.cmask = (int32_t)groups.size();
                    groups.push_back(gd);
                    launches.back().stats[du.f == 5 ? 3 : 2] += 1;
                } else {
                    launches.back().stats[1] += 1;
                }
                ops.push_back(o);
            }
            pd.ngroups = 0;
            passes.push_back(pd);
            continue;
        }
        // groups: maximal runs of ops whose targets fit R register bits; a dense op is a group of its own
        size_t i = 0;
        double log2_growth = 0;   // bound on the growth of the tile's largest |amplitude| over the pass
        // diagonal table: 2^k complex64 entries (16-byte aligned) appended to the operand array;
        // returns their float offset; |entries| bound the growth
        auto put_table = [&](const DenseUnit &du) {
            while (wmat.size() % 4) wmat.push_back(0.f);
            const int64_t off = (int64_t)wmat.size();
            double amax = 0;
            for (size_t x = 0; 2 * x < du.u.size(); ++x) {
                wmat.push_back((float)du.u[2 * x]);
                wmat.push_back((float)du.u[2 * x + 1]);
                amax = std::max(amax, std::hypot(du.u[2 * x], du.u[2 * x + 1]));
            }
            log2_growth += std::log2(std::max(1.0, amax));
            return off;
        };
        // the same table inside a MUX2 register group (register bits rb, in register order): entry
        // r 2^(K-R) + s belongs to register r of slot s (element: s with zero bits inserted at the
        // register bits, ascending, then register r's bits set), so the threads of a wavefront read
        // consecutive entries for each register
        auto put_table_regs = [&](const DenseUnit &du, const std::vector<int> &rb) {
            while (wmat.size() % 4) wmat.push_back(0.f);
            const int64_t off = (int64_t)wmat.size();
            std::vector<int> bs(rb.begin(), rb.end());
            std::sort(bs.begin(), bs.end());
            const uint32_t nslot = 1u << (K - R), kmask = (1u << k) - 1u;
            wmat.resize(wmat.size() + (size_t(2) << K));
            double amax = 0;
            for (uint32_t sl = 0; sl < nslot; ++sl) {
                uint32_t base = sl;
                for (int q = 0; q < R; ++q) base = ((base >> bs[q]) << (bs[q] + 1)) | (base & ((1u << bs[q]) - 1u));
                for (uint32_t r = 0; r < (1u << R); ++r) {
                    uint32_t e = base;
                    for (int q = 0; q < R; ++q)
                        if ((r >> q) & 1u) e |= 1u << rb[q];
                    const size_t x = e & kmask, o = off + 2 * (size_t(r) * nslot + sl);
                    wmat[o] = (float)du.u[2 * x];
                    wmat[o + 1] = (float)du.u[2 * x + 1];
                    amax = std::max(amax, std::hypot(du.u[2 * x], du.u[2 * x + 1]));
                }
            }
            log2_growth += std::log2(std::max(1.0, amax));
            return off;
        };
        auto is_table = [&](size_t x) { return dense_of[x] >= 0 && dus[dense_of[x]].diag; };
        // one dense unit as a group (FP32: three-product or 32x32x2 A operands; split FP16: hi / lo)
        auto add_dense = [&](size_t i) -> int {
                const DenseUnit &du = dus[dense_of[i]];
                const int f0 = du.f, F = std::max(4, f0);
                GroupDev gd{};
                gd.f = F;
                std::vector<int> qs(du.q, du.q + f0);
                for (int b = 0; (int)qs.size() < F && b < K; ++b)
                    if (std::find(qs.begin(), qs.end(), b) == qs.end()) qs.push_back(b);
                if ((int)qs.size() < F) return fail(SVX_E_CAPACITY, "tile too small for a dense op");
                // tile bit 0 inside the unit becomes the matrix's index bit 0 (the 16-byte LDS path of
                // the dense-only kernel pairs components 2c, 2c + 1): swap index bits 0 and p0
                int p0 = 0;
                for (int q = 0; q < F; ++q)
                    if (qs[q] == 0) p0 = q;
                std::swap(qs[0], qs[p0]);
                auto sw = [p0](int x) {
                    const int b0 = x & 1, bp = (x >> p0) & 1;
                    return (x & ~1 & ~(1 << p0)) | (bp) | (b0 << p0);
                };
                for (int q = 0; q < F; ++q) gd.dq[q] = gd.dqs[q] = qs[q];
                std::sort(gd.dqs, gd.dqs + F);
                // padded matrix: the extra qubits are the high index bits (identity on them), in the
                // index order before the bit-0 swap
                const int D = 1 << F, d0 = 1 << f0;
                auto U = [&](int r, int c, int part) -> double {
                    r = sw(r);
                    c = sw(c);
                    if ((r >> f0) != (c >> f0)) return 0.0;
                    return du.u[2 * ((r & (d0 - 1)) * d0 + (c & (d0 - 1))) + part];
                };
                // real form W = [[Ur, -Ui], [Ui, Ur]] (2D x 2D) acting on [Re x; Im x]
                auto W = [&](int row, int col) -> double {
                    const int r = row % D, c = col % D;
                    const bool rim = row >= D, cim = col >= D;
                    if (rim == cim) return U(r, c, 0);
                    return rim ? U(r, c, 1) : -U(r, c, 1);
                };
                // entries in {0, +-1}: the FP32 products are exact (permutations stay bit-exact)
                bool exact = true;
                for (double v : du.u) exact = exact && (v == 0.0 || v == 1.0 || v == -1.0);
                const int NRT = F == 5 ? 2 : 1;
                while (wmat.size() % 4) wmat.push_back(0.f);   // 16-byte aligned operand blocks
                gd.w_off = (int64_t)wmat.size();
                gd.exact = exact ? 1 : 0;
                if (dense_f32 && !exact && F == 4) {
                    // FP32, three real products (Gauss): wd[(s 4 + t) 64 + l] = W_s[l % 16][4 (l / 16) + t],
                    // W_0 = Ur + Ui, W_1 = -Ui, W_2 = Ur (A operands of v_mfma_f32_16x16x4f32)
                    gd.type = 5;
                    gd.jmask = (int32_t)unit_lane_bits(K, (1u << qs[0]) | (1u << qs[1]) | (1u << qs[2]) | (1u << qs[3]));
                    for (int sidx = 0; sidx < 3; ++sidx)
                        for (int t = 0; t < 4; ++t)
                            for (int l = 0; l < 64; ++l) {
                                const int r = l % 16, c = 4 * (l / 16) + t;
                                const double ur = U(r, c, 0), ui = U(r, c, 1);
                                wmat.push_back((float)(sidx == 0 ? ur + ui : sidx == 1 ? -ui : ur));
                            }
                } else if (dense_f32 || exact) {
                    // FP32 A operand of v_mfma_f32_32x32x2f32: wd[(rt 2^F + s) 64 + l] = W[32 rt + l % 32][2 s + l / 32]
                    gd.type = 1;
                    for (int rt = 0; rt < NRT; ++rt)
                        for (int st = 0; st < D; ++st)
                            for (int l = 0; l < 64; ++l) wmat.push_back((float)W(32 * rt + l % 32, 2 * st + l / 32));
                } else {
                    // split-FP16 A operands of v_mfma_f32_32x32x16_f16: 2^wexp W = Wh + Wl with the
                    // largest |entry| of 2^wexp W below 2^14; block ((rt NKS + ks) 2 + part) holds
                    // lane l, entry e = Wi[32 rt + l % 32][16 ks + 8 (l / 32) + e], Wi the interleaved
                    // real form (index 2c + p: component c, part p) -- Wi[r][s] = W[il(r)][il(s)]
                    gd.type = 3;
                    auto il = [D](int x) { return (x & 1) * D + (x >> 1); };
                    const int NKS = D / 8;
                    double wmax = 0;
                    for (double v : du.u) wmax = std::max(wmax, std::fabs(v));
                    int wE = 0;
                    if (wmax > 0) std::frexp(wmax, &wE);   // wmax < 2^wE
                    // entries within [2^-4, 2]: no scaling (the low halves' FP16 spacing 2^-24 stays
                    // below FP32 rounding of the products); otherwise the largest entry goes below 2^14
                    gd.wexp = (wmax == 0 || (wmax >= 0.0625 && wmax <= 2.0)) ? 0 : 14 - wE;
                    const double wsc = std::ldexp(1.0, gd.wexp);
                    std::vector<uint16_t> hb;
                    hb.reserve((size_t)NRT * NKS * 2 * 512);
                    for (int rt = 0; rt < NRT; ++rt)
                        for (int ks = 0; ks < NKS; ++ks)
                            for (int part = 0; part < 2; ++part)
                                for (int l = 0; l < 64; ++l)
                                    for (int e = 0; e < 8; ++e) {
                                        const double ws = W(il(32 * rt + l % 32), il(16 * ks + 8 * (l / 32) + e)) * wsc;
                                        const uint16_t hh = f32_to_f16_bits((float)ws);
                                        hb.push_back(part == 0 ? hh : f32_to_f16_bits((float)(ws - f16_bits_to_f32(hh))));
                                    }
                    const size_t o = wmat.size();
                    wmat.resize(o + hb.size() / 2);
                    std::memcpy(wmat.data() + o, hb.data(), hb.size() * sizeof(uint16_t));
                }
                // max-abs growth bound of the op: largest row sum of |U| (padding adds identity rows)
                double rs_max = 0;
                for (int r = 0; r < d0; ++r) {
                    double rs = 0;
                    for (int c = 0; c < d0; ++c) rs += std::hypot(du.u[2 * (r * d0 + c)], du.u[2 * (r * d0 + c) + 1]);
                    rs_max = std::max(rs_max, rs);
                }
                log2_growth += std::log2(std::max(1.0, rs_max));
                groups.push_back(gd);
                return 0;
        };
        if (use_mux) {
            // dense units in place; each run of MUX2 ops and tables between them as register groups
            size_t a0 = 0;
            while (a0 < pops.size()) {
                if (dense_of[a0] >= 0 && !is_table(a0)) {
                    const int rcd = add_dense(a0);
                    if (rcd) return rcd;
                    ++a0;
                    continue;
                }
                size_t b0 = a0;
                while (b0 < pops.size() && !(dense_of[b0] >= 0 && !is_table(b0))) ++b0;
                const std::vector<OpDev<T>> sub(pops.begin() + a0, pops.begin() + b0);
                const std::vector<int> sub_dense(dense_of.begin() + a0, dense_of.begin() + b0);
                const int rc = mux_groups<T>(sub, sub_dense, dus, K, R, opt.mux == SVX_MUX_AUTO, groups, ops,
                                             put_table_regs, log2_growth, &launches.back().stats[5]);
                if (rc) return rc;
                a0 = b0;
            }
            i = pops.size();
        }
        while (i < pops.size()) {
            if (is_table(i)) {
                // a table with no MUX2 op before the next dense unit (or the end): its own group, the
                // LDS tile swept in order (contiguous table reads); beside MUX2 ops it joins their group
                size_t j = i;
                while (j < pops.size() && is_table(j)) ++j;
                if (j == pops.size() || dense_of[j] >= 0) {
                    GroupDev gd{};
                    gd.type = 6;
                    gd.w_off = put_table(dus[dense_of[i]]);
                    groups.push_back(gd);
                    ++i;
                    continue;
                }
            }
            if (dense_of[i] >= 0 && !is_table(i)) {
                const int rcd = add_dense(i);
                if (rcd) return rcd;
                ++i;
                continue;
            }
            std::vector<int> tg;
            size_t j = i;
            for (; j < pops.size(); ++j) {
                if (is_table(j)) continue;   // a table needs no register bits
                if (dense_of[j] >= 0) break;
                const int t = pops[j].t;
                if (std::find(tg.begin(), tg.end(), t) == tg.end()) {
                    if ((int)tg.size() == R) break;
                    tg.push_back(t);
                }
            }
            // register bits: the targets, then filler bits from the top of the workgroup tile
            std::vector<int> rb = tg;
            for (int b = K - 1; b >= 0 && (int)rb.size() < R; --b)
                if (std::find(rb.begin(), rb.end(), b) == rb.end()) rb.push_back(b);
            GroupDev gd{};
            gd.nops = (int)(j - i);
            gd.op0 = (int)ops.size();
            uint32_t regmask = 0;
            for (int q = 0; q < R; ++q) {
                gd.b[q] = gd.bs[q] = rb[q];
                regmask |= 1u << rb[q];
            }
            std::sort(gd.bs, gd.bs + R);
            for (size_t o = i; o < j; ++o) {
                if (is_table(o)) {   // each register's amplitude times its table entry
                    OpDev<T> od{};
                    od.kind = K_TABLE;
                    od.aux = (int32_t)put_table_regs(dus[dense_of[o]], rb);
                    ops.push_back(od);
                    continue;
                }
                OpDev<T> od = pops[o];
                od.tr = (int)(std::find(rb.begin(), rb.end(), od.t) - rb.begin());
                uint32_t cmr = 0;
                for (int q = 0; q < R; ++q)
                    if ((uint32_t)od.cmask >> rb[q] & 1u) cmr |= 1u << q;
                od.cmr = cmr;
                od.cmt = (uint32_t)od.cmask & ~regmask;
                if (od.cmask == 0 || od.kind <= K_XSWAP) {
                    od.cmr = 0;
                    od.cmt = 0;
                }
                // one control, inside the register group: the touched pairs are known at compile time
                if (od.cmt == 0 && __builtin_popcount(od.cmr) == 1) {
                    if (od.kind == K_CXSWAP) od.kind = K_CXSWAP1;
                    else if (od.kind == K_CPHASE) od.kind = K_CPHASE1;
                }
                ops.push_back(od);
                auto rows = [](const T *mm) {
                    const double a = std::hypot((double)mm[0], (double)mm[1]) + std::hypot((double)mm[2], (double)mm[3]);
                    const double b = std::hypot((double)mm[4], (double)mm[5]) + std::hypot((double)mm[6], (double)mm[7]);
                    return std::max(a, b);
                };
                log2_growth += std::log2(std::max(1.0, std::max(rows(od.m0), rows(od.m1))));
            }
            groups.push_back(gd);
            i = j;
        }
        // Tile scaling for the split-FP16 groups: the kernel scales each tile so its largest
        // component is below 2^tscale; 2^tscale times the pass's growth bound (and sqrt 2 for
        // re/im against complex magnitudes) stays below 2^15, the FP16 range the split needs.
        bool split16 = false;   // split-FP16 dense groups (the only ones that need the scaled tile)
        for (int g = pd.group0; g < (int)groups.size(); ++g) split16 = split16 || groups[g].type == 3;
        if (mode == 1 && split16 && tile_scale_ok) {
            const int h = (int)std::ceil(log2_growth + 0.5 + 0.02);
            if (15 - h >= 4) {
                pd.tscale = 15 - h;
                for (int g = pd.group0; g < (int)groups.size(); ++g)
                    if (groups[g].type == 3) groups[g].type = 4;
            }
        }
        pd.ngroups = (int)groups.size() - pd.group0;
        LaunchInfo &lr = launches.back();
        // dense-only passes of 4-qubit units run on the 512-thread dense-only kernel (dense_only = 0:
        // never); with FP32 units the MUX2 kernel runs them (measured slightly faster: 194.7 against
        // 197.1 ms for the 30-qubit workload), so only split-FP16 passes take it
        if (mode == 1 && has_dense && opt.dense_only && !use_mux && (K == 12 || K == 13) && pd.ngroups > 0 &&
            k - rp.m >= 1) {
            bool all4 = true;
            for (int g = pd.group0; g < (int)groups.size(); ++g) all4 = all4 && groups[g].type != 0 && groups[g].type != 6 && groups[g].f == 4;
            lr.d4 = all4;
            if (all4) lr.block = kNTD4;
        }
        for (int g = pd.group0; g < (int)groups.size(); ++g) {
            const GroupDev &G = groups[g];
            if (G.type == 0) {
                int nt = 0;
                for (int o = 0; o < G.nops; ++o)
                    nt += lr.mux ? ops[G.op0 + o].kind == (M_TABLE - M_FULL) * 8 : ops[G.op0 + o].kind == K_TABLE;
                lr.stats[0] += 1;
                lr.stats[1] += G.nops - nt;
                lr.stats[6] += nt;
            } else if (G.type == 6) {
                lr.stats[6] += 1;
            } else {
                lr.stats[G.f == 5 ? 3 : 2] += 1;
                lr.stats[4] += G.exact;
            }
        }
        passes.push_back(pd);
    }
    // layout: passes, groups, ops (256-byte aligned sections)
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    // the same passes with the identity workgroup -> tile order, for svx_tune_prepared's choice
    std::vector<PassDev> alt = passes;
    for (PassDev &a : alt) a.torder = 0;
    pr->pass_off = 0;
    pr->alt_off = al(passes.size() * sizeof(PassDev));
    pr->group_off = pr->alt_off + al(alt.size() * sizeof(PassDev));
    pr->op_off = pr->group_off + al(groups.size() * sizeof(GroupDev));
    pr->w_off = pr->op_off + al(ops.size() * sizeof(OpDev<T>));
    const size_t total = std::max<size_t>(256, pr->w_off + al(wmat.size() * sizeof(float)));
    blob->assign(total, 0);
    if (!passes.empty()) std::memcpy(blob->data() + pr->pass_off, passes.data(), passes.size() * sizeof(PassDev));
    if (!alt.empty()) std::memcpy(blob->data() + pr->alt_off, alt.data(), alt.size() * sizeof(PassDev));
    pr->alt.reset(new std::atomic<uint8_t>[passes.size() + 1]);
    for (size_t i = 0; i <= passes.size(); ++i) pr->alt[i].store(0);
    pr->torder = torder;
    if (!groups.empty()) std::memcpy(blob->data() + pr->group_off, groups.data(), groups.size() * sizeof(GroupDev));
    if (!ops.empty()) std::memcpy(blob->data() + pr->op_off, ops.data(), ops.size() * sizeof(OpDev<T>));
    if (!wmat.empty()) std::memcpy(blob->data() + pr->w_off, wmat.data(), wmat.size() * sizeof(float));
    pr->launches = launches;
    return 0;
}

// Free the descriptors of unreferenced entries, oldest release first, while their bytes exceed the
// cache limit; an entry whose last launch has not completed is kept for a later call.
void evict_locked(DeviceState *ds) {
    if (g_unref_bytes <= g_unref_limit) return;
    std::vector<std::pair<uint64_t, int64_t>> cand;
    for (size_t i = 0; i < g_prepared.size(); ++i) {
        const Prepared *p = g_prepared[i].get();
        if (p && p->refs == 0 && p->device == ds->caps.device) cand.push_back({p->lru, (int64_t)i + 1});
    }
    std::sort(cand.begin(), cand.end());
    for (const auto &c : cand) {
        if (g_unref_bytes <= g_unref_limit) break;
        Prepared *p = g_prepared[c.second - 1].get();
        if (p->last_use && hipEventQuery(p->last_use) != hipSuccess) continue;   // still in flight
        auto range = g_prep_index.equal_range(p->key);
        for (auto it = range.first; it != range.second; ++it)
            if (it->second == c.second) {
                g_prep_index.erase(it);
                break;
            }
        ds->arena.release(p->dev, p->dev_bytes);
        g_unref_bytes -= p->dev_bytes;
        if (p->last_use) (void)hipEventDestroy(p->last_use);
        g_prepared[c.second - 1].reset();
    }
}

int64_t find_locked(uint64_t key, const int32_t *w, int64_t nw, int dtype, int nbits, int device, const Opts &o) {
    auto range = g_prep_index.equal_range(key);
    for (auto it = range.first; it != range.second; ++it) {
        Prepared *p = g_prepared[it->second - 1].get();
        if (p && p->dtype == dtype && p->nbits == nbits && p->device == device &&
            std::memcmp(&p->opts, &o, sizeof(Opts)) == 0 && (int64_t)p->words.size() == nw &&
            std::memcmp(p->words.data(), w, nw * 4) == 0)
            return it->second;
    }
    return 0;
}

void addref_locked(int64_t h) {
    Prepared *p = g_prepared[h - 1].get();
    if (p->refs++ == 0) g_unref_bytes -= p->dev_bytes;
}

// Look up or build the prepared form of a segment; the caller receives one reference.  The
// build (classification, fusion search, operand layout) runs outside the registry lock.
int prepare(const int32_t *w, int64_t nw, int dtype, int nbits, const svx_options *opts, int64_t *handle,
            Prepared **out) {
    if (!w || nw <= 0) return fail(SVX_E_ARG, "empty QSV2 stream");
    if (dtype != SVX_C64 && dtype != SVX_C128) return fail(SVX_E_ARG, "dtype must be SVX_C64 or SVX_C128");
    Opts o;
    int rc = normalize_opts(opts, &o);
    if (rc) return rc;
    DeviceState *ds = nullptr;
    rc = get_ds(&ds);
    if (rc) return rc;
    const uint64_t h = fnv1a(w, nw, (uint64_t)dtype * 1315423911u + (uint64_t)nbits * 2654435761u +
                                        (uint64_t)ds->caps.device * 97u +
                                        fnv1a(reinterpret_cast<const int32_t *>(&o), sizeof(Opts) / 4, 7));
    {
        std::lock_guard<std::mutex> lk(g_prep_mu);
        if (int64_t hd = find_locked(h, w, nw, dtype, nbits, ds->caps.device, o)) {
            addref_locked(hd);
            *handle = hd;
            if (out) *out = g_prepared[hd - 1].get();
            return 0;
        }
    }
    RawSeg seg;
    rc = parse_segment(w, nw, &seg);
    if (rc) return rc;
    if (seg.L > nbits) return fail(SVX_E_ARG, "segment L exceeds the state size");
    std::unique_ptr<Prepared> pr(new Prepared());
    pr->dtype = dtype;
    pr->nbits = nbits;
    pr->device = ds->caps.device;
    pr->words.assign(w, w + nw);
    pr->opts = o;
    pr->key = h;
    std::vector<unsigned char> blob;
    rc = dtype == SVX_C64 ? build_prepared<float>(seg, nbits, dtype, o, pr.get(), &blob)
                          : build_prepared<double>(seg, nbits, dtype, o, pr.get(), &blob);
    if (rc) return rc;
    void *dev = nullptr;
    size_t got = 0;
    {
        std::lock_guard<std::mutex> lk(g_prep_mu);
        if (int64_t hd = find_locked(h, w, nw, dtype, nbits, ds->caps.device, o)) {   // built meanwhile
            addref_locked(hd);
            *handle = hd;
            if (out) *out = g_prepared[hd - 1].get();
            return 0;
        }
        evict_locked(ds);
        rc = ds->arena.alloc(blob.size(), &dev, &got);
        if (rc) return rc;
    }
    hipError_t e = hipMemcpy(dev, blob.data(), blob.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pr->last_use, hipEventDisableTiming);
    std::lock_guard<std::mutex> lk(g_prep_mu);
    if (e != hipSuccess) {
        ds->arena.release(dev, got);
        if (pr->last_use) (void)hipEventDestroy(pr->last_use);
        return hipfail(e, "descriptor upload");
    }
    pr->dev = dev;
    pr->dev_bytes = got;
    pr->refs = 1;
    g_prepared.push_back(std::move(pr));
    const int64_t hd = (int64_t)g_prepared.size();
    g_prep_index.insert({h, hd});
    *handle = hd;
    if (out) *out = g_prepared.back().get();
    return 0;
}

int release(int64_t handle) {
    std::lock_guard<std::mutex> lk(g_prep_mu);
    Prepared *p = lookup_locked(handle);
    if (!p) return fail(SVX_E_ARG, "unknown or released prepared-segment handle");
    if (--p->refs == 0) {
        g_unref_bytes += p->dev_bytes;
        p->lru = ++g_lru_clock;
        DeviceState *ds = nullptr;
        std::lock_guard<std::mutex> lk2(g_mu);
        auto it = g_dev.find(p->device);
        if (it != g_dev.end()) ds = it->second.get();
        if (ds) evict_locked(ds);
    }
    return 0;
}

// After a launch sequence that read a prepared segment's descriptors: remember its completion
// (not under stream capture: a captured graph's owner keeps its handles referenced).
void mark_use(Prepared *pr, hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    (void)hipEventRecord(pr->last_use, s);
}

// one pass launch of a prepared segment (default or tuned tile order)
hipError_t launch_one(void *state, const Prepared *pr, const LaunchInfo &li, hipStream_t s) {
    unsigned char *base = static_cast<unsigned char *>(pr->dev);
    const bool alt = pr->alt[li.pass_index].load(std::memory_order_relaxed) != 0;
    const PassDev *pd = reinterpret_cast<const PassDev *>(base + (alt ? pr->alt_off : pr->pass_off)) + li.pass_index;
    const GroupDev *gd = reinterpret_cast<const GroupDev *>(base + pr->group_off);
    const float *wd = li.dense ? reinterpret_cast<const float *>(base + pr->w_off) : nullptr;
    if (li.d4)
        return launch_pass_d4(li.K, state, pd, gd, reinterpret_cast<const OpDev<float> *>(base + pr->op_off), wd,
                              li.grid, li.lds, s);
    if (li.mux)
        return launch_mux(li.K, li.units, state, pd, gd, reinterpret_cast<const OpDev<float> *>(base + pr->op_off), wd,
                          li.grid, li.lds, s);
    if (pr->dtype == SVX_C64)
        return launch_pass_c64(li.K, li.small, state, pd, gd, reinterpret_cast<const OpDev<float> *>(base + pr->op_off),
                               wd, li.grid, li.lds, s);
    return launch_pass_c128(li.K, li.small, state, pd, gd, reinterpret_cast<const OpDev<double> *>(base + pr->op_off),
                            li.grid, li.lds, s);
}

int run_prepared(void *state, Prepared *pr, hipStream_t s) {
    if (!state) return fail(SVX_E_ARG, "null state pointer");
    if (reinterpret_cast<uintptr_t>(state) & 15) return fail(SVX_E_ARG, "state must be 16-byte aligned");
    for (const LaunchInfo &li : pr->launches) {
        if (li.empty) continue;
        const hipError_t e = launch_one(state, pr, li, s);
        if (e != hipSuccess) return hipfail(e, "pass kernel launch");
    }
    mark_use(pr, s);
    return 0;
}

// A referenced handle (error otherwise); the caller's reference keeps it alive after the lock.
int get_prepared(int64_t handle, Prepared **out) {
    std::lock_guard<std::mutex> lk(g_prep_mu);
    Prepared *p = lookup_locked(handle);
    if (!p) return fail(SVX_E_ARG, "unknown or released prepared-segment handle");
    *out = p;
    return 0;
}

// ------------------------------------------------------------------------------------------ permute
// Permutation descriptors live in a bounded cache keyed by [device, n, perm...].  Each entry records
// an event after its launches; an entry is evicted (least recently used first) only once that event
// has completed, and entries used under stream capture are kept (a captured graph reads them later).
struct PermEntry {
    void *dev = nullptr;
    uint32_t grid = 0;
    hipEvent_t ev = nullptr;
    uint64_t lru = 0;
    bool pinned = false;
};
constexpr size_t kPermCacheMax = 64;
std::mutex g_perm_mu;
std::map<std::vector<int32_t>, PermEntry> g_perm_cache;
uint64_t g_perm_tick = 0;

// Tile of a permutation sweep: v = min(n, 12) bits holding new bits 0..c-1 (c = min(6, n)) and the new
// positions of old bits 0..c-1, so reads and writes both run over 2^c contiguous amplitudes; the
// other n - v bits are fixed per workgroup.  `tile` (optional) names further bits the tile must hold:
// an in-place step passes the bits it moves, and since the step maps that set onto itself and fixes
// every other bit, each workgroup reads and writes the same 2^v amplitudes.
int build_perm(int n, const int32_t *perm, PermDev *pd, uint32_t *grid, const std::vector<int> *tile = nullptr) {
    std::vector<int> inv(n, -1);
    for (int b = 0; b < n; ++b) {
        if (perm[b] < 0 || perm[b] >= n || inv[perm[b]] != -1) return fail(SVX_E_ARG, "perm is not a permutation of 0..n-1");
        inv[perm[b]] = b;
    }
    std::memset(pd, 0, sizeof(*pd));
    int c, v;
    if (n <= 12) {
        c = std::min(6, n);
        v = n;
    } else {
        c = 6;
        v = 12;
    }
    std::vector<char> inV(n, 0);
    for (int b = 0; b < c; ++b) inV[b] = 1;
    for (int o = 0; o < c; ++o) inV[inv[o]] = 1;
    if (tile)
        for (int b : *tile) inV[b] = 1;
    int cnt = 0;
    for (int b = 0; b < n; ++b) cnt += inV[b];
    for (int b = 0; b < n && cnt < v; ++b)
        if (!inV[b]) {
            inV[b] = 1;
            ++cnt;
        }
    std::vector<int> V, F;
    for (int b = 0; b < n; ++b) (inV[b] ? V : F).push_back(b);
    if ((int)V.size() != v) return fail(SVX_E_ARG, "internal: permutation tile size");
    std::vector<int> rank(n, -1);
    for (int i = 0; i < v; ++i) rank[V[i]] = i;
    std::vector<int> OV;   // old positions of the tile bits, ascending
    for (int i = 0; i < v; ++i) OV.push_back(perm[V[i]]);
    std::sort(OV.begin(), OV.end());
    for (int p = 0; p < c; ++p)
        if (OV[p] != p) return fail(SVX_E_ARG, "internal: permutation low bits");
    pd->n = n;
    pd->c = c;
    pd->v = v;
    pd->nfix = (int)F.size();
    for (size_t i = 0; i < F.size(); ++i) {
        pd->fixnew[i] = F[i];
        pd->fixold[i] = perm[F[i]];
    }
    // read-order index q: bit p <-> old bit OV[p] -> new bit inv[OV[p]] -> local bit rank[...]
    auto rdpos = [&](uint32_t q) {
        uint32_t pos = 0;
        for (int p = 0; p < v; ++p)
            if ((q >> p) & 1u) pos |= 1u << rank[inv[OV[p]]];
        return pos;
    };
    for (uint32_t x = 0; x < 64; ++x) {
        pd->rdlo[x] = (uint16_t)(x < (1u << std::min(v, 6)) ? rdpos(x) : 0);
        pd->rdhi[x] = (uint16_t)((v > 6 && x < (1u << (v - 6))) ? rdpos(x << 6) : 0);
        uint64_t oh = 0, nh = 0;
        if (x < (1u << (v - c))) {
            for (int p = c; p < v; ++p)
                if ((x >> (p - c)) & 1u) {
                    oh |= uint64_t(1) << OV[p];
                    nh |= uint64_t(1) << V[p];
                }
        }
        pd->oldhi[x] = oh;
        pd->newhi[x] = nh;
    }
    *grid = (uint32_t)(uint64_t(1) << (n - v));
    return 0;
}

// Bin configuration of the FP64 marginal reduction (k_marg) and its grid.
int marg_cfg(int n, const int32_t *fmap, int n_low, MargCfg *cfg, uint32_t *grid) {
    std::memset(cfg, 0, sizeof(*cfg));
    const int c = std::min(n, 13);
    cfg->n = n;
    cfg->c = c;
    cfg->low8 = std::min(c, 8);
    cfg->nlow = n_low;
    int used[64];
    for (int b = 0; b < n; ++b) used[b] = -1;
    for (int i = 0; i < n_low; ++i) {
        if (fmap[i] < 0 || fmap[i] >= n || used[fmap[i]] != -1) return fail(SVX_E_ARG, "fmap entries must be distinct bits < n");
        used[fmap[i]] = i;
    }
    for (int b = 0; b < cfg->low8; ++b) {
        if (used[b] >= 0) {
            cfg->sa[cfg->nsa] = b;
            cfg->sa_x[cfg->nsa++] = used[b];
        } else {
            cfg->ra[cfg->nra++] = b;
        }
    }
    for (int b = cfg->low8; b < c; ++b) {
        if (used[b] >= 0) {
            cfg->sb[cfg->nsb] = b;
            cfg->sb_x[cfg->nsb++] = used[b];
        } else {
            cfg->rb[cfg->nrb++] = b;
        }
    }
    for (int b = c; b < n; ++b)
        if (used[b] >= 0) {
            cfg->shi[cfg->nshi] = b - c;
            cfg->shi_x[cfg->nshi++] = used[b];
        }
    const uint64_t nchunk = uint64_t(1) << (n - c);
    *grid = (uint32_t)std::min<uint64_t>(nchunk, 2048);
    return 0;
}

}  // namespace

// ============================================================================================ C ABI
extern "C" {

const char *svx_last_error(void) { return g_err.c_str(); }
const char *svx_version(void) { return "libsvx 0.1 (gfx950)"; }

int svx_init(int32_t device, svx_caps *out) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(SVX_E_NODEV, "no HIP device is visible");
    if (device < 0 || device >= n) return fail(SVX_E_NODEV, "no HIP device " + std::to_string(device));
    DeviceState *ds = nullptr;
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != device) HIPCHK(hipSetDevice(device));
    int rc = ensure_device(device, &ds);
    if (cur != device) (void)hipSetDevice(cur);
    if (rc) return rc;
    if (out) *out = ds->caps;
    return 0;
}

int svx_default_options(svx_options *out) {
    if (!out) return fail(SVX_E_ARG, "null options pointer");
    default_opts(out);
    return 0;
}

int svx_prepare_ex(const int32_t *qsv2, int64_t n_words, int32_t dtype, int32_t state_bits, const svx_options *opts,
                   int64_t *handle) {
    if (!handle) return fail(SVX_E_ARG, "null handle pointer");
    if (state_bits < 1 || state_bits > 40) return fail(SVX_E_ARG, "state_bits out of range");
    return prepare(qsv2, n_words, dtype, state_bits, opts, handle, nullptr);
}

int svx_prepare(const int32_t *qsv2, int64_t n_words, int32_t dtype, int32_t state_bits, int64_t *handle) {
    return svx_prepare_ex(qsv2, n_words, dtype, state_bits, nullptr, handle);
}

int svx_release_prepared(int64_t handle) { return release(handle); }

int svx_cache_limit(int64_t limit_bytes, int64_t *held) {
    if (limit_bytes < 0) return fail(SVX_E_ARG, "negative cache limit");
    DeviceState *ds = nullptr;
    int rc = get_ds(&ds);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_prep_mu);
    g_unref_limit = (size_t)limit_bytes;
    evict_locked(ds);
    if (held) {
        size_t b = 0;
        for (const auto &p : g_prepared)
            if (p && p->device == ds->caps.device) b += p->dev_bytes;
        *held = (int64_t)b;
    }
    return 0;
}

int svx_prepared_shape(int64_t handle, int32_t *dtype, int32_t *state_bits) {
    Prepared *pr = nullptr;
    int rc = get_prepared(handle, &pr);
    if (rc) return rc;
    if (dtype) *dtype = pr->dtype;
    if (state_bits) *state_bits = pr->nbits;
    return 0;
}

int svx_run_prepared(void *state, int64_t handle, void *stream) {
    Prepared *pr = nullptr;
    int rc = get_prepared(handle, &pr);
    if (rc) return rc;
    return run_prepared(state, pr, (hipStream_t)stream);
}

int svx_tune_prepared(void *state, int64_t handle, int32_t reps, void *stream) {
    Prepared *pr = nullptr;
    int rc = get_prepared(handle, &pr);
    if (rc) return rc;
    if (!state) return fail(SVX_E_ARG, "null state pointer");
    if (reinterpret_cast<uintptr_t>(state) & 15) return fail(SVX_E_ARG, "state must be 16-byte aligned");
    if (reps < 1 || reps > 64) return fail(SVX_E_ARG, "reps must be in 1..64");
    const hipStream_t s = (hipStream_t)stream;
    hipEvent_t ev[2];
    HIPCHK(hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return hipfail(hipErrorOutOfMemory, "event");
    }
    for (const LaunchInfo &li : pr->launches) {
        if (li.empty || !li.tunable) continue;
        float ms[2] = {0.f, 0.f};
        for (int c = 0; c < 2 && !rc; ++c) {
            pr->alt[li.pass_index].store((uint8_t)c);
            hipError_t e = launch_one(state, pr, li, s);   // untimed: first launch of this order
            if (e == hipSuccess) e = hipEventRecord(ev[0], s);
            for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_one(state, pr, li, s);
            if (e == hipSuccess) e = hipEventRecord(ev[1], s);
            if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms[c], ev[0], ev[1]);
            if (e != hipSuccess) rc = hipfail(e, "tile-order timing");
        }
        // the identity order only when clearly faster (timing noise stays on the default)
        pr->alt[li.pass_index].store((!rc && ms[1] < 0.97f * ms[0]) ? 1 : 0);
        if (rc) break;
    }
    mark_use(pr, s);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return rc;
}

int svx_run_batch(void *const *states, const int64_t *handles, const int32_t *passes, int32_t n_items,
                  void *stream) {
    if (n_items < 0 || (n_items > 0 && (!states || !handles || !passes))) return fail(SVX_E_ARG, "bad batch arrays");
    if (n_items == 0) return 0;
    if (n_items > DeviceState::kBatchCap) return fail(SVX_E_ARG, "more than 4096 items in one batch");
    DeviceState *ds = nullptr;
    int rc = get_ds(&ds);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    struct Cls {
        std::vector<BatchItem> items;
        uint64_t tiles = 0;
        size_t lds = 0;
    };
    std::map<int, Cls> cls;   // key: K * 8 + dense + 2 d4 + 4 mux
    std::vector<Prepared *> used;
    {
        std::set<void *> seen;
        for (int i = 0; i < n_items; ++i)
            if (!seen.insert(states[i]).second)
                return fail(SVX_E_ARG, "batch item " + std::to_string(i) + " repeats a state of an earlier item");
        std::lock_guard<std::mutex> lk(g_prep_mu);
        for (int i = 0; i < n_items; ++i) {
            Prepared *pr = lookup_locked(handles[i]);
            if (!pr) return fail(SVX_E_ARG, "unknown or released prepared-segment handle in batch item " + std::to_string(i));
            if (pr->dtype != SVX_C64) return fail(SVX_E_ARG, "batched launches run complex64 segments");
            if (pr->device != ds->caps.device) return fail(SVX_E_ARG, "batch item prepared on another device");
            if (passes[i] < 0 || passes[i] >= (int32_t)pr->launches.size())
                return fail(SVX_E_ARG, "bad pass index in batch item " + std::to_string(i));
            if (!states[i] || (reinterpret_cast<uintptr_t>(states[i]) & 15))
                return fail(SVX_E_ARG, "batch states must be non-null and 16-byte aligned");
            const LaunchInfo &li = pr->launches[passes[i]];
            if (li.empty) continue;
            if (li.small)
                return fail(SVX_E_ARG, "batch item " + std::to_string(i) +
                                           ": only passes of the single-buffer kernels batch (state >= 2^10 amplitudes)");
            unsigned char *base = static_cast<unsigned char *>(pr->dev);
            BatchItem it{};
            it.st = states[i];
            it.pd = reinterpret_cast<const PassDev *>(base + pr->pass_off) + li.pass_index;
            it.gd = reinterpret_cast<const GroupDev *>(base + pr->group_off);
            it.od = base + pr->op_off;
            it.wd = li.dense ? reinterpret_cast<const float *>(base + pr->w_off) : nullptr;
            Cls &c = cls[li.K * 16 + (li.dense ? 1 : 0) + (li.d4 ? 2 : 0) + (li.mux ? 4 : 0) + (li.units ? 8 : 0)];
            it.tile0 = c.tiles;
            c.tiles += li.ntiles;
            c.lds = std::max(c.lds, li.lds);
            c.items.push_back(it);
            used.push_back(pr);
        }
    }
    std::lock_guard<std::mutex> lk(ds->mu);
    if (!ds->batch_host) {
        HIPCHK(hipHostMalloc((void **)&ds->batch_host, sizeof(BatchItem) * DeviceState::kBatchCap * DeviceState::kBatchSlots));
        HIPCHK(hipMalloc((void **)&ds->batch_dev, sizeof(BatchItem) * DeviceState::kBatchCap * DeviceState::kBatchSlots));
        for (int i = 0; i < DeviceState::kBatchSlots; ++i) HIPCHK(hipEventCreateWithFlags(&ds->batch_ev[i], hipEventDisableTiming));
    }
    for (auto &kv : cls) {
        Cls &c = kv.second;
        const int bK = kv.first / 16;
        const bool bmux = (kv.first & 4) != 0;
        const uint64_t threads = bmux ? (uint64_t(1) << (bK - mux_register_bits(bK))) : ((kv.first & 2) ? kNTD4 : kNT);
        if (c.tiles * threads > 0xffffffffull)
            return fail(SVX_E_CAPACITY, "batch launch beyond the launch limit of 2^32 work-items");
        const int slot = ds->batch_next;
        ds->batch_next = (ds->batch_next + 1) % DeviceState::kBatchSlots;
        if (ds->batch_used[slot]) HIPCHK(hipEventSynchronize(ds->batch_ev[slot]));
        BatchItem *hs = ds->batch_host + (size_t)slot * DeviceState::kBatchCap;
        BatchItem *dv = ds->batch_dev + (size_t)slot * DeviceState::kBatchCap;
        std::memcpy(hs, c.items.data(), c.items.size() * sizeof(BatchItem));
        HIPCHK(hipMemcpyAsync(dv, hs, c.items.size() * sizeof(BatchItem), hipMemcpyHostToDevice, s));
        const hipError_t e = bmux ? launch_mux_batch(bK, (kv.first & 8) != 0, dv, (int)c.items.size(), (uint32_t)c.tiles, c.lds, s)
                                  : launch_pass_batch(bK, kv.first & 1, (kv.first & 2) != 0, dv, (int)c.items.size(),
                                                      (uint32_t)c.tiles, c.lds, s);
        if (e != hipSuccess) return hipfail(e, "batched pass launch");
        HIPCHK(hipEventRecord(ds->batch_ev[slot], s));
        ds->batch_used[slot] = true;
    }
    std::sort(used.begin(), used.end());
    used.erase(std::unique(used.begin(), used.end()), used.end());
    for (Prepared *pr : used) mark_use(pr, s);
    return 0;
}

int svx_prepared_info(int64_t handle, int32_t *n_passes, int32_t *n_launches) {
    Prepared *pr = nullptr;
    int rc = get_prepared(handle, &pr);
    if (rc) return rc;
    int nl = 0;
    for (auto &li : pr->launches) nl += li.empty ? 0 : 1;
    if (n_passes) *n_passes = (int32_t)pr->launches.size();
    if (n_launches) *n_launches = nl;
    return 0;
}

int svx_prepared_kernel(int64_t handle, int32_t pass, char *name, int32_t cap) {
    Prepared *pr = nullptr;
    int rc = get_prepared(handle, &pr);
    if (rc) return rc;
    if (pass < 0 || pass >= (int32_t)pr->launches.size() || !name || cap < 1) return fail(SVX_E_ARG, "bad pass index");
    const LaunchInfo &li = pr->launches[pass];
    const char *T = pr->dtype == SVX_C64 ? "float" : "double";
    std::string k;
    if (li.empty) k = "none";
    else if (li.small) k = std::string("k_pass_small<") + T + ">";
    else if (li.d4) k = "k_pass_d4<" + std::to_string(li.K) + ">";
    else if (li.mux) k = "k_mux<" + std::to_string(li.K) + "," + std::to_string(mux_register_bits(li.K)) + ">";
    else k = std::string("k_pass<") + T + "," + std::to_string(li.K) + (li.dense ? ",true>" : ",false>");
    std::snprintf(name, cap, "%s", k.c_str());
    return 0;
}

int svx_prepared_pass_stats(int64_t handle, int32_t pass, int32_t *out) {
    Prepared *pr = nullptr;
    int rc = get_prepared(handle, &pr);
    if (rc) return rc;
    if (pass < 0 || pass >= (int32_t)pr->launches.size() || !out) return fail(SVX_E_ARG, "bad pass index");
    for (int i = 0; i < 7; ++i) out[i] = pr->launches[pass].stats[i];
    out[7] = pr->alt[pr->launches[pass].pass_index].load() ? 0 : pr->torder;
    return 0;
}

int svx_run_segment_ex(void *state, int64_t n_slice_amps, int32_t n_slices, int32_t dtype, const int32_t *qsv2,
                       int64_t n_words, const svx_options *opts, void *stream) {
    const int L = ilog2_exact(n_slice_amps);
    if (L < 1) return fail(SVX_E_ARG, "n_slice_amps must be a power of two >= 2");
    if (n_slices < 1) return fail(SVX_E_ARG, "n_slices must be >= 1");
    if (!qsv2 || n_words < kHdr) return fail(SVX_E_FORMAT, "QSV2 stream shorter than its 8-word header");
    if ((uint32_t)qsv2[0] != kMagic) return fail(SVX_E_MAGIC, "bad QSV2 magic");
    if (qsv2[1] != L)
        return fail(SVX_E_ARG, "segment L = " + std::to_string(qsv2[1]) + " does not match the slice of 2^" + std::to_string(L));
    const int pb = ilog2_exact(n_slices);
    Prepared *pr = nullptr;
    int64_t h = 0;
    // a power-of-two slice count: one launch sequence over all slices; otherwise one per slice
    int rc = prepare(qsv2, n_words, dtype, pb >= 0 ? L + pb : L, opts, &h, &pr);
    if (rc) return rc;
    if (pb >= 0) {
        rc = run_prepared(state, pr, (hipStream_t)stream);
    } else {
        const size_t es = dtype == SVX_C64 ? 8 : 16;
        for (int32_t p = 0; p < n_slices && !rc; ++p)
            rc = run_prepared(static_cast<unsigned char *>(state) + (size_t)p * n_slice_amps * es, pr, (hipStream_t)stream);
    }
    const std::string err = g_err;
    (void)release(h);
    if (rc) g_err = err;
    return rc;
}

int svx_run_segment(void *state, int64_t n_slice_amps, int32_t n_slices, int32_t dtype, const int32_t *qsv2,
                    int64_t n_words, void *stream) {
    return svx_run_segment_ex(state, n_slice_amps, n_slices, dtype, qsv2, n_words, nullptr, stream);
}

int svx_engine(void *st, int64_t ins_off, void *stream) {
    if (!st || ins_off < 0) return fail(SVX_E_ARG, "bad buffer or instruction offset");
    hipStream_t s = (hipStream_t)stream;
    const unsigned char *ins = static_cast<const unsigned char *>(st) + ins_off * 32;
    int32_t hdr[kHdr];
    HIPCHK(hipMemcpyAsync(hdr, ins, sizeof(hdr), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if ((uint32_t)hdr[0] != kMagic) return fail(SVX_E_MAGIC, "bad QSV2 magic");
    // walk pass headers to find the stream length
    std::vector<int32_t> words(hdr, hdr + kHdr);
    int64_t pos = kHdr;
    for (int p = 0; p < hdr[2]; ++p) {
        int32_t ph[kPassW];
        HIPCHK(hipMemcpy(ph, ins + pos * 4, sizeof(ph), hipMemcpyDeviceToHost));
        int64_t len = kPassW + (ph[0] > 4 ? kPassW : 0);
        for (int o = 0; o < std::max(0, ph[3]); ++o) {
            int32_t oh[2];
            HIPCHK(hipMemcpy(oh, ins + (pos + len) * 4, sizeof(oh), hipMemcpyDeviceToHost));
            len += oh[0] == kOpDense ? 8 + (int64_t(2) << (2 * std::min(5, std::max(1, oh[1])))) : kOpW;
            if (len > (int64_t(1) << 20)) return fail(SVX_E_FORMAT, "pass longer than the 4 MiB instruction region");
        }
        const size_t old = words.size();
        words.resize(old + len);
        HIPCHK(hipMemcpy(words.data() + old, ins + pos * 4, len * 4, hipMemcpyDeviceToHost));
        pos += len;
    }
    const int L = hdr[1];
    if (ins_off * 32 < (int64_t(1) << L) * 8) return fail(SVX_E_ARG, "instruction region overlaps the 2^L slice");
    return svx_run_segment(st, int64_t(1) << L, 1, SVX_C64, words.data(), (int64_t)words.size(), stream);
}

// In-place plan of perm[new] = old: steps s_1 .. s_S, each a permutation that moves only bits of a set
// M_i with at most 6 bits outside the low 6 (so M_i fits one 12-bit sweep tile, `tile`), applied in
// order (position b after step i holds what position s_i[b] held before it).  Position b needs the
// bit now at need(b); the cycles of `need` are taken whole while the tile has room (smallest first),
// else as a segment b_1 .. b_j of one cycle, which places j - 1 bits and leaves the cycle j - 1
// shorter.  Each step is one in-place sweep; at most n steps.
static int plan_inplace(int n, const int32_t *perm, std::vector<std::vector<int32_t>> *steps,
                        std::vector<std::vector<int>> *tiles) {
    const int c = std::min(6, n), budget = std::min(12, n) - c;
    std::vector<int> cur(n), pos(n);   // cur[b]: original bit at position b; pos: its inverse
    for (int b = 0; b < n; ++b) cur[b] = pos[b] = b;
    for (int guard = 0; guard <= n; ++guard) {
        std::vector<int> need(n);
        bool done = true;
        for (int b = 0; b < n; ++b) {
            need[b] = pos[perm[b]];
            done = done && need[b] == b;
        }
        if (done) return 0;
        // cycles b -> need(b) -> need(need(b)) ... of the positions not yet in place
        std::vector<char> seen(n, 0);
        std::vector<std::vector<int>> cyc;
        for (int b = 0; b < n; ++b) {
            if (seen[b] || need[b] == b) continue;
            std::vector<int> cy;
            for (int x = b; !seen[x]; x = need[x]) {
                seen[x] = 1;
                cy.push_back(x);
            }
            cyc.push_back(cy);
        }
        auto high = [&](const std::vector<int> &v) {
            int h = 0;
            for (int x : v) h += x >= c;
            return h;
        };
        std::sort(cyc.begin(), cyc.end(), [&](const std::vector<int> &a, const std::vector<int> &b) {
            return high(a) < high(b);
        });
        std::vector<char> inM(n, 0);
        int used = 0;
        bool any = false;
        for (const auto &cy : cyc) {   // whole cycles that fit
            const int h = high(cy);
            if (used + h > budget) continue;
            for (int x : cy) inM[x] = 1;
            used += h;
            any = true;
        }
        if (!any) {   // the longest segment of the first cycle within the budget
            const auto &cy = cyc[0];
            const int L = (int)cy.size();
            int best_len = 0, best_start = 0;
            for (int st = 0; st < L; ++st) {
                int h = 0, len = 0;
                while (len < L) {
                    const int hx = cy[(st + len) % L] >= c;
                    if (h + hx > budget) break;
                    h += hx;
                    ++len;
                }
                if (len > best_len) {
                    best_len = len;
                    best_start = st;
                }
            }
            if (best_len < 2) return fail(SVX_E_ARG, "internal: in-place permutation plan");
            for (int t = 0; t < best_len; ++t) inM[cy[(best_start + t) % L]] = 1;
        }
        // the step: b in M takes need(b) when that lies in M; the rest of M takes the unclaimed bits
        std::vector<int32_t> stp(n);
        std::vector<int> open, spare;
        std::vector<char> claimed(n, 0);
        for (int b = 0; b < n; ++b) {
            stp[b] = b;
            if (!inM[b]) continue;
            if (inM[need[b]]) {
                stp[b] = need[b];
                claimed[need[b]] = 1;
            } else {
                open.push_back(b);
            }
        }
        for (int b = 0; b < n; ++b)
            if (inM[b] && !claimed[b]) spare.push_back(b);
        if (open.size() != spare.size()) return fail(SVX_E_ARG, "internal: in-place permutation step");
        for (size_t t = 0; t < open.size(); ++t) stp[open[t]] = spare[t];
        std::vector<int> tile;
        for (int b = 0; b < n; ++b)
            if (inM[b]) tile.push_back(b);
        std::vector<int> nc(n);
        for (int b = 0; b < n; ++b) nc[b] = cur[stp[b]];
        cur = nc;
        for (int b = 0; b < n; ++b) pos[cur[b]] = b;
        steps->push_back(stp);
        tiles->push_back(tile);
    }
    return fail(SVX_E_ARG, "internal: in-place permutation did not converge");
}

// A cached permutation descriptor (built and uploaded on first use), evicting the least recently
// used completed entry when the cache is full.
static int perm_entry(DeviceState *ds, int n, const int32_t *perm, const std::vector<int> *tile, hipStream_t s,
                      PermEntry **out) {
    std::vector<int32_t> key{ds->caps.device, n, tile ? 1 : 0};
    key.insert(key.end(), perm, perm + n);
    auto it = g_perm_cache.find(key);
    if (it == g_perm_cache.end()) {
        PermDev pd;
        uint32_t grid = 0;
        int rc = build_perm(n, perm, &pd, &grid, tile);
        if (rc) return rc;
        while (g_perm_cache.size() >= kPermCacheMax) {
            auto victim = g_perm_cache.end();
            for (auto jt = g_perm_cache.begin(); jt != g_perm_cache.end(); ++jt)
                if (!jt->second.pinned && jt->first[0] == ds->caps.device &&
                    (victim == g_perm_cache.end() || jt->second.lru < victim->second.lru))
                    victim = jt;
            if (victim == g_perm_cache.end()) break;   // every entry is pinned by a captured graph
            HIPCHK(hipEventSynchronize(victim->second.ev));
            HIPCHK(hipFree(victim->second.dev));
            (void)hipEventDestroy(victim->second.ev);
            g_perm_cache.erase(victim);
        }
        PermEntry e;
        e.grid = grid;
        HIPCHK(hipMalloc(&e.dev, sizeof(PermDev)));
        HIPCHK(hipMemcpy(e.dev, &pd, sizeof(PermDev), hipMemcpyHostToDevice));
        HIPCHK(hipEventCreateWithFlags(&e.ev, hipEventDisableTiming));
        it = g_perm_cache.emplace(key, e).first;
    }
    it->second.lru = ++g_perm_tick;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) it->second.pinned = true;
    *out = &it->second;
    return 0;
}

static int perm_launch(int dtype, void *dst, const void *src, int n, PermEntry *pe, hipStream_t s) {
    const int v = n <= 12 ? n : 12;
    // the tile (2 padding elements per 32) and the descriptor's four 64-entry tables
    const size_t lds = ((size_t(1) << v) + ((size_t(1) << v) >> 4)) * (dtype == SVX_C64 ? 8 : 16) + 64 * (8 + 8 + 2 + 2);
    hipError_t e = launch_permute(dtype, dst, src, (const PermDev *)pe->dev, pe->grid, lds, s);
    if (e != hipSuccess) return hipfail(e, "permute launch");
    if (!pe->pinned) (void)hipEventRecord(pe->ev, s);
    return 0;
}

int svx_permute(void *dst, const void *src, int32_t n, const int32_t *perm, int32_t dtype, void *stream) {
    if (!src || !perm) return fail(SVX_E_ARG, "null pointer");
    if (!dst) dst = const_cast<void *>(src);   // NULL destination: in place
    if (n < 1 || n > 40) return fail(SVX_E_ARG, "n out of range");
    if (dtype != SVX_C64 && dtype != SVX_C128) return fail(SVX_E_ARG, "bad dtype");
    const size_t bytes = (size_t(1) << n) * (dtype == SVX_C64 ? 8 : 16);
    const char *pd = static_cast<const char *>(dst), *ps = static_cast<const char *>(src);
    if (dst != src && pd < ps + bytes && ps < pd + bytes) return fail(SVX_E_ARG, "dst and src overlap partially");
    {
        std::vector<int> seen(n, 0);
        for (int b = 0; b < n; ++b) {
            if (perm[b] < 0 || perm[b] >= n || seen[perm[b]]) return fail(SVX_E_ARG, "perm is not a permutation of 0..n-1");
            seen[perm[b]] = 1;
        }
    }
    DeviceState *ds = nullptr;
    int rc = get_ds(&ds);
    if (rc) return rc;
    const hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_perm_mu);
    if (dst != src) {
        PermEntry *pe = nullptr;
        rc = perm_entry(ds, n, perm, nullptr, s, &pe);
        return rc ? rc : perm_launch(dtype, dst, src, n, pe, s);
    }
    std::vector<std::vector<int32_t>> steps;
    std::vector<std::vector<int>> tiles;
    rc = plan_inplace(n, perm, &steps, &tiles);
    if (rc) return rc;
    for (size_t i = 0; i < steps.size(); ++i) {
        PermEntry *pe = nullptr;
        rc = perm_entry(ds, n, steps[i].data(), &tiles[i], s, &pe);
        if (rc) return rc;
        rc = perm_launch(dtype, dst, dst, n, pe, s);
        if (rc) return rc;
    }
    return 0;
}

int svx_permute_plan(int32_t n, const int32_t *perm, int32_t in_place, int32_t *steps, int32_t max_steps,
                     int32_t *n_steps) {
    if (!perm || !n_steps || n < 1 || n > 40 || max_steps < 0) return fail(SVX_E_ARG, "bad permute-plan arguments");
    std::vector<int> seen(n, 0);
    for (int b = 0; b < n; ++b) {
        if (perm[b] < 0 || perm[b] >= n || seen[perm[b]]) return fail(SVX_E_ARG, "perm is not a permutation of 0..n-1");
        seen[perm[b]] = 1;
    }
    std::vector<std::vector<int32_t>> st;
    std::vector<std::vector<int>> tiles;
    if (in_place) {
        const int rc = plan_inplace(n, perm, &st, &tiles);
        if (rc) return rc;
    } else {
        st.emplace_back(perm, perm + n);
    }
    *n_steps = (int32_t)st.size();
    if (steps) {
        if ((int32_t)st.size() > max_steps) return fail(SVX_E_ARG, "steps buffer too small");
        for (size_t i = 0; i < st.size(); ++i) std::memcpy(steps + i * n, st[i].data(), n * sizeof(int32_t));
    }
    return 0;
}

int svx_swap_words(void *a, void *b, int64_t offa, int64_t offb, int64_t nwords, void *stream) {
    if (!a || !b || offa < 0 || offb < 0 || nwords < 0) return fail(SVX_E_ARG, "bad swap arguments");
    const char *pa = static_cast<char *>(a) + offa * 64, *pb = static_cast<char *>(b) + offb * 64;
    if (pa < pb + nwords * 64 && pb < pa + nwords * 64 && nwords > 0) return fail(SVX_E_ARG, "swap ranges overlap");
    hipError_t e = launch_swap_words(a, b, offa, offb, nwords, (hipStream_t)stream);
    if (e != hipSuccess) return hipfail(e, "swap launch");
    return 0;
}

int svx_zero_state(void *state, int32_t n, int32_t dtype, void *stream) {
    if (!state || n < 0 || n > 40) return fail(SVX_E_ARG, "bad zero-state arguments");
    if (dtype != SVX_C64 && dtype != SVX_C128) return fail(SVX_E_ARG, "bad dtype");
    hipError_t e = launch_zero(dtype, state, int64_t(1) << n, (hipStream_t)stream);
    if (e != hipSuccess) return hipfail(e, "zero launch");
    return 0;
}

int svx_convert(void *dst, int32_t dt_dst, const void *src, int32_t dt_src, int64_t n_amps, void *stream) {
    if (!dst || !src || n_amps < 0) return fail(SVX_E_ARG, "bad convert arguments");
    hipError_t e = launch_convert(dst, dt_dst, src, dt_src, n_amps, (hipStream_t)stream);
    if (e != hipSuccess) return hipfail(e, "convert launch");
    return 0;
}

static int ensure_scratch(DeviceState *ds, size_t part_bytes, size_t misc_bytes) {
    if (ds->part_bytes < part_bytes) {
        if (ds->part) HIPCHK(hipFree(ds->part));
        ds->part = nullptr;
        HIPCHK(hipMalloc(&ds->part, part_bytes));
        ds->part_bytes = part_bytes;
    }
    if (ds->misc_bytes < misc_bytes) {
        if (ds->misc) HIPCHK(hipFree(ds->misc));
        ds->misc = nullptr;
        HIPCHK(hipMalloc(&ds->misc, misc_bytes));
        ds->misc_bytes = misc_bytes;
    }
    return 0;
}

int svx_product_state(void *state, int32_t n, int32_t dtype, const double *amps, void *stream) {
    if (!state || !amps || n < 1 || n > 40) return fail(SVX_E_ARG, "bad product-state arguments");
    if (dtype != SVX_C64 && dtype != SVX_C128) return fail(SVX_E_ARG, "bad dtype");
    DeviceState *ds = nullptr;
    int rc = get_ds(&ds);
    if (rc) return rc;
    const int nlo = n / 2, nhi = n - nlo;
    auto table = [&](int q0, int nq) {
        std::vector<double> t(size_t(2) << nq);
        for (uint64_t x = 0; x < (uint64_t(1) << nq); ++x) {
            double re = 1, im = 0;
            for (int i = 0; i < nq; ++i) {
                const double *a = amps + 4 * (q0 + i) + (((x >> i) & 1) ? 2 : 0);
                const double r2 = re * a[0] - im * a[1], i2 = re * a[1] + im * a[0];
                re = r2;
                im = i2;
            }
            t[2 * x] = re;
            t[2 * x + 1] = im;
        }
        return t;
    };
    std::vector<double> tlo = table(0, nlo), thi = table(nlo, nhi);
    std::lock_guard<std::mutex> lk(ds->mu);
    const size_t bytes = (tlo.size() + thi.size()) * sizeof(double) + 256;
    rc = ensure_scratch(ds, std::max<size_t>(ds->part_bytes, 1 << 20), std::max<size_t>(bytes, 1 << 16));
    if (rc) return rc;
    double *dlo = static_cast<double *>(ds->misc);
    double *dhi = dlo + tlo.size();
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(dlo, tlo.data(), tlo.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dhi, thi.data(), thi.size() * sizeof(double), hipMemcpyHostToDevice));
    hipError_t e = launch_product(dtype, state, n, dlo, dhi, nlo, s);
    if (e != hipSuccess) return hipfail(e, "product launch");
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int svx_reduce_f64(const void *state, int32_t n, int32_t dtype, const int32_t *fmap, int32_t n_low, double *out_norm,
                   double *out_marginals, void *stream) {
    if (!state || n < 1 || n > 40 || n_low < 0 || n_low > 8 || n_low > n) return fail(SVX_E_ARG, "bad reduce arguments");
    if (n_low > 0 && (!fmap || !out_marginals)) return fail(SVX_E_ARG, "fmap and out_marginals required when n_low > 0");
    if (dtype != SVX_C64 && dtype != SVX_C128) return fail(SVX_E_ARG, "bad dtype");
    DeviceState *ds = nullptr;
    int rc = get_ds(&ds);
    if (rc) return rc;
    MargCfg cfg;
    uint32_t grid = 0;
    rc = marg_cfg(n, fmap, n_low, &cfg, &grid);
    if (rc) return rc;
    const int nbin = 1 << n_low;
    std::lock_guard<std::mutex> lk(ds->mu);
    rc = ensure_scratch(ds, std::max<size_t>((size_t)grid * 256 * sizeof(double), 1 << 20), 1 << 16);
    if (rc) return rc;
    unsigned char *misc = static_cast<unsigned char *>(ds->misc);
    MargCfg *cfg_dev = reinterpret_cast<MargCfg *>(misc);
    double *out_dev = reinterpret_cast<double *>(misc + 4096);
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(hipMemcpy(cfg_dev, &cfg, sizeof(cfg), hipMemcpyHostToDevice));
    hipError_t e = launch_marginals(dtype, state, cfg_dev, nbin, ds->part, grid, out_dev, s);
    if (e != hipSuccess) return hipfail(e, "marginal launch");
    double out[256];
    HIPCHK(hipMemcpyAsync(out, out_dev, nbin * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double norm = 0;
    for (int x = 0; x < nbin; ++x) norm += out[x];
    if (out_norm) *out_norm = norm;
    if (n_low > 0) std::memcpy(out_marginals, out, nbin * sizeof(double));
    return 0;
}

// svx_reduce_f64 without allocation: the FP64 partials, the bin configuration and the folded bins
// live in this device's copy of the file-scope `scratch` array (its address looked up at device
// set-up).  Same kernels, grid and reduction order as svx_reduce_f64, so the results are identical.
int svx_reduce_f64_noalloc(const void *state, int32_t n, int32_t dtype, const int32_t *fmap, int32_t n_low,
                           double *out_norm, double *out_marginals, void *stream) {
    if (!state || n < 1 || n > 40 || n_low < 0 || n_low > n) return fail(SVX_E_ARG, "bad reduce arguments");
    if (n_low > kScratchMaxLow)
        return fail(SVX_E_CAPACITY, "svx_reduce_f64_noalloc: 2^" + std::to_string(n_low) + " bins exceed its " +
                                        std::to_string(kScratchBytes) + "-byte scratch array (at most 2^" +
                                        std::to_string(kScratchMaxLow) + " bins of 2048 workgroups)");
    if (n_low > 0 && (!fmap || !out_marginals)) return fail(SVX_E_ARG, "fmap and out_marginals required when n_low > 0");
    if (dtype != SVX_C64 && dtype != SVX_C128) return fail(SVX_E_ARG, "bad dtype");
    DeviceState *ds = nullptr;
    int rc = get_ds(&ds);
    if (rc) return rc;
    if (!ds->scratch) return fail(SVX_E_HIP, "scratch array address unavailable on this device");
    MargCfg cfg;
    uint32_t grid = 0;
    rc = marg_cfg(n, fmap, n_low, &cfg, &grid);
    if (rc) return rc;
    const int nbin = 1 << n_low;
    if ((size_t)grid * nbin * sizeof(double) > kScratchPartBytes)
        return fail(SVX_E_CAPACITY, "svx_reduce_f64_noalloc: partials exceed the scratch array");
    // layout of the scratch array: [partials: kScratchPartBytes][MargCfg: 4 KiB][bins: 256 doubles]
    double *part = reinterpret_cast<double *>(ds->scratch);
    MargCfg *cfg_dev = reinterpret_cast<MargCfg *>(ds->scratch + kScratchPartBytes);
    double *out_dev = reinterpret_cast<double *>(ds->scratch + kScratchPartBytes + 4096);
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(hipMemcpyAsync(cfg_dev, &cfg, sizeof(cfg), hipMemcpyHostToDevice, s));
    hipError_t e = launch_marginals(dtype, state, cfg_dev, nbin, part, grid, out_dev, s);
    if (e != hipSuccess) return hipfail(e, "marginal launch");
    double out[1 << kScratchMaxLow];
    HIPCHK(hipMemcpyAsync(out, out_dev, nbin * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double norm = 0;
    for (int x = 0; x < nbin; ++x) norm += out[x];
    if (out_norm) *out_norm = norm;
    if (n_low > 0) std::memcpy(out_marginals, out, nbin * sizeof(double));
    return 0;
}

int svx_fidelity_f64(const void *a, int32_t dta, const void *b, int32_t dtb, int32_t n, double *out, void *stream) {
    if (!a || !b || !out || n < 0 || n > 40) return fail(SVX_E_ARG, "bad fidelity arguments");
    DeviceState *ds = nullptr;
    int rc = get_ds(&ds);
    if (rc) return rc;
    const int64_t N = int64_t(1) << n;
    const uint32_t grid = (uint32_t)std::max<int64_t>(1, std::min<int64_t>(2048, N / 1024));
    std::lock_guard<std::mutex> lk(ds->mu);
    rc = ensure_scratch(ds, std::max<size_t>((size_t)grid * 4 * sizeof(double), 1 << 20), 1 << 16);
    if (rc) return rc;
    unsigned char *misc = static_cast<unsigned char *>(ds->misc);
    double *o4 = reinterpret_cast<double *>(misc + 8192);
    double *coef = o4 + 8;
    double *o2 = coef + 8;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = launch_fid_pass1(dta, a, dtb, b, N, ds->part, grid, o4, s);
    if (e != hipSuccess) return hipfail(e, "fidelity launch");
    double h4[4];
    HIPCHK(hipMemcpyAsync(h4, o4, sizeof(h4), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double na = std::sqrt(h4[2]), nb = std::sqrt(h4[3]);
    if (!(na > 0) || !(nb > 0)) return fail(SVX_E_ARG, "zero-norm state in fidelity");
    const double ipr = h4[0] / (na * nb), ipi = h4[1] / (na * nb);
    const double ab = std::sqrt(ipr * ipr + ipi * ipi);
    double hc[4] = {ab > 0 ? ipr / ab : 1.0, ab > 0 ? ipi / ab : 0.0, 1.0 / na, 1.0 / nb};
    // e = b/|b| - ph a/|a| with ph = <a|b>/|<a|b>| minimises |e|
    HIPCHK(hipMemcpy(coef, hc, sizeof(hc), hipMemcpyHostToDevice));
    e = launch_fid_pass2(dta, a, dtb, b, N, coef, ds->part, grid, o2, s);
    if (e != hipSuccess) return hipfail(e, "fidelity launch");
    double h2[2];
    HIPCHK(hipMemcpyAsync(h2, o2, sizeof(h2), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    out[0] = ab * ab;
    out[1] = h2[0] - h2[0] * h2[0] / 4;
    out[2] = std::sqrt(h2[1]);
    out[3] = std::sqrt(h2[0]);
    return 0;
}

int svx_sync(void *stream) {
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

// ------------------------------------------------------------------------------------------ queue
// Persistent queue of whole-state circuits (k_mux_queue).  Jobs and completion words live in
// host-visible coherent memory; the claim counter in device memory.  One resident launch serves
// every job published while it runs; it drains by itself when idle, and submit / done / wait
// relaunch it when jobs are pending and the previous launch has finished.
struct SmallQueue {
    int device = 0, K = 0;
    uint32_t cap = 0, wgs = 0;
    size_t lds = 0;
    uint64_t idle_ticks = 0;
    hipStream_t stream = nullptr;
    QueueJob *jobs = nullptr;
    QueueCtl *ctl = nullptr;
    uint32_t *next = nullptr;
    uint64_t tickets = 0;
    bool launched = false;
    std::mutex mu;
};
std::mutex g_queue_mu;
std::vector<std::unique_ptr<SmallQueue>> g_queues;

static int get_queue(int64_t q, SmallQueue **out) {
    std::lock_guard<std::mutex> lk(g_queue_mu);
    if (q < 1 || q > (int64_t)g_queues.size() || !g_queues[q - 1]) return fail(SVX_E_ARG, "unknown or closed queue");
    *out = g_queues[q - 1].get();
    return 0;
}

// (Re)launch unless a launch is resident or every job is done (caller holds q->mu).  Jobs are
// claimed in ticket order and a finished launch completed every job it claimed, so after a finished
// launch the newest job is done exactly when all are.
static int queue_kick(SmallQueue *q) {
    if (q->tickets == 0) return 0;
    if (q->launched) {
        const hipError_t st = hipStreamQuery(q->stream);
        if (st == hipErrorNotReady) return 0;   // still resident
        if (st != hipSuccess) return hipfail(st, "queue kernel");
        const uint64_t last = q->tickets - 1;
        if (__atomic_load_n(&q->ctl->done[last % q->cap], __ATOMIC_ACQUIRE) == (uint32_t)(last + 1)) return 0;
    }
    const hipError_t e = launch_mux_queue(q->K, q->jobs, q->ctl, q->next, q->cap, q->idle_ticks, q->wgs, q->lds, q->stream);
    if (e != hipSuccess) return hipfail(e, "queue kernel launch");
    q->launched = true;
    return 0;
}

int svx_queue_open(int32_t state_bits, int32_t capacity, int32_t workgroups, int32_t idle_us, int64_t *queue) {
    if (!queue || state_bits < 10 || state_bits > 13 || capacity < 1 || capacity > (1 << 20) || workgroups < 1 ||
        workgroups > 4096 || idle_us < 1)
        return fail(SVX_E_ARG, "bad queue arguments (state_bits 10..13, capacity 1..2^20, workgroups 1..4096)");
    DeviceState *ds = nullptr;
    int rc = get_ds(&ds);
    if (rc) return rc;
    auto q = std::make_unique<SmallQueue>();
    q->device = ds->caps.device;
    q->K = state_bits;
    q->cap = (uint32_t)capacity;
    q->wgs = (uint32_t)workgroups;
    q->idle_ticks = uint64_t(idle_us) * 100;   // 100 MHz real-time counter
    const size_t ne = size_t(1) << state_bits;
    q->lds = (ne + (ne >> 5)) * 8 + 128 * 8 + 64;
    HIPCHK(hipHostMalloc((void **)&q->jobs, sizeof(QueueJob) * q->cap, hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(hipHostMalloc((void **)&q->ctl, sizeof(QueueCtl) + sizeof(uint32_t) * q->cap,
                         hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(q->jobs, 0, sizeof(QueueJob) * q->cap);
    std::memset(q->ctl, 0, sizeof(QueueCtl) + sizeof(uint32_t) * q->cap);
    HIPCHK(hipMalloc((void **)&q->next, sizeof(uint32_t)));
    HIPCHK(hipMemset(q->next, 0, sizeof(uint32_t)));
    HIPCHK(hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking));
    std::lock_guard<std::mutex> lk(g_queue_mu);
    g_queues.push_back(std::move(q));
    *queue = (int64_t)g_queues.size();
    return 0;
}

int svx_queue_submit(int64_t queue, void *state, int64_t handle, int32_t zero_init, int64_t *ticket) {
    SmallQueue *q = nullptr;
    int rc = get_queue(queue, &q);
    if (rc) return rc;
    if (!state || (reinterpret_cast<uintptr_t>(state) & 15)) return fail(SVX_E_ARG, "state must be non-null and 16-byte aligned");
    Prepared *pr = nullptr;
    rc = get_prepared(handle, &pr);
    if (rc) return rc;
    if (pr->dtype != SVX_C64 || pr->nbits != q->K || pr->device != q->device)
        return fail(SVX_E_ARG, "queue jobs are complex64 states of the queue's size on its device");
    if (pr->launches.size() != 1 || !pr->launches[0].mux || pr->launches[0].ntiles != 1)
        return fail(SVX_E_ARG, "queue jobs are one whole-state pass on the MUX2 kernel (compile with k = state bits)");
    const LaunchInfo &li = pr->launches[0];
    std::lock_guard<std::mutex> lk(q->mu);
    const uint64_t t = q->tickets;
    const uint32_t slot = (uint32_t)(t % q->cap);
    if (t >= q->cap && __atomic_load_n(&q->ctl->done[slot], __ATOMIC_ACQUIRE) != (uint32_t)(t - q->cap + 1))
        return fail(SVX_E_CAPACITY, "queue full: every slot holds a job that is not done");
    unsigned char *base = static_cast<unsigned char *>(pr->dev);
    QueueJob &jb = q->jobs[slot];
    jb.st = reinterpret_cast<uint64_t>(state);
    jb.pd = reinterpret_cast<uint64_t>(reinterpret_cast<const PassDev *>(base + pr->pass_off) + li.pass_index);
    jb.gd = reinterpret_cast<uint64_t>(base + pr->group_off);
    jb.od = reinterpret_cast<uint64_t>(base + pr->op_off);
    jb.wd = reinterpret_cast<uint64_t>(base + pr->w_off);
    jb.flags = zero_init ? 1 : 0;
    q->tickets = t + 1;
    __atomic_store_n(&q->ctl->tail, (uint32_t)(t + 1), __ATOMIC_RELEASE);
    rc = queue_kick(q);
    if (rc) return rc;
    *ticket = (int64_t)t;
    return 0;
}

int svx_queue_done(int64_t queue, int64_t ticket, int32_t *done) {
    SmallQueue *q = nullptr;
    int rc = get_queue(queue, &q);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(q->mu);
    if (ticket < 0 || (uint64_t)ticket >= q->tickets) return fail(SVX_E_ARG, "unknown ticket");
    const uint32_t w = __atomic_load_n(&q->ctl->done[ticket % q->cap], __ATOMIC_ACQUIRE);
    if ((uint64_t)ticket + q->cap < q->tickets) {   // its slot was reused: done long ago
        *done = 1;
        return 0;
    }
    *done = w == (uint32_t)(ticket + 1);
    if (!*done) return queue_kick(q);
    return 0;
}

int svx_queue_wait(int64_t queue, int64_t ticket) {
    for (int64_t spin = 0;; ++spin) {
        int32_t d = 0;
        const int rc = svx_queue_done(queue, ticket, &d);
        if (rc || d) return rc;
        if (spin > 4000000) return fail(SVX_E_HIP, "queue job did not complete");
        std::this_thread::sleep_for(std::chrono::microseconds(spin < 1000 ? 1 : 20));
    }
}

int svx_queue_close(int64_t queue) {
    std::unique_ptr<SmallQueue> q;
    {
        std::lock_guard<std::mutex> lk(g_queue_mu);
        if (queue < 1 || queue > (int64_t)g_queues.size() || !g_queues[queue - 1]) return fail(SVX_E_ARG, "unknown or closed queue");
        q = std::move(g_queues[queue - 1]);
    }
    __atomic_store_n(&q->ctl->stop, 1u, __ATOMIC_RELEASE);
    HIPCHK(hipStreamSynchronize(q->stream));
    HIPCHK(hipStreamDestroy(q->stream));
    HIPCHK(hipFree(q->next));
    HIPCHK(hipHostFree(q->jobs));
    HIPCHK(hipHostFree(q->ctl));
    return 0;
}

}  // extern "C"
